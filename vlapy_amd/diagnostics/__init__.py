"""Host-side diagnostics of the field history (vlapy/diagnostics)."""
