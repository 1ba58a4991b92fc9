"""MI355X-native drop-in for the RF hot paths of GeRaF (geraf.ops, geraf.models.model_utils)."""
