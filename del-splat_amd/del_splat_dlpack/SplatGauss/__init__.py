"""Gaussian splatting pipeline (torch front end in SplatGauss.torch)."""
