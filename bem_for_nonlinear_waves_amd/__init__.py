"""MI355X-native FMM-accelerated GMRES BVP solver (BEM-MEL Laplace BIE)."""
