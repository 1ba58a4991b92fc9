"""Per-step RF operations of the stage renderers (calibration, MF prediction and loss)."""
