"""Backend-dispatch helpers between the renderers and the RF ops."""
