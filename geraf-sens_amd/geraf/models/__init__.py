"""Model-side helpers of the RF path (only model_utils is provided)."""
