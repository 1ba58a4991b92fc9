"""Producer networks of the RF path (SDF, variance, reflective power)."""
