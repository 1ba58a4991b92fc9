"""RF operators backed by the gfx950 C-ABI library libgeraf_rf.so."""
