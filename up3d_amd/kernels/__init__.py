"""Device element kernels (pyfp3d/kernels)."""
