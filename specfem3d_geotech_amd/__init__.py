"""MI355X-native EBE-PCG + viscoplastic hot path for SPECFEM3D_GEOTECH."""
