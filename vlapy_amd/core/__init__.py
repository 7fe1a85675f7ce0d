"""vlapy.core call surface (vlasov, field, collisions, vlasov_poisson, step)
backed by the HIP kernels of libvpfp.so."""
