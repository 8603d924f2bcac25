"""Learned absorbed-power surrogate for the coupled epoch loop (device network,
tile features, training on GPU-MC data, judge/fallback coupling)."""
