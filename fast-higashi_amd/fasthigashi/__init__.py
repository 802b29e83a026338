"""MI355X-native Fast-Higashi decomposition path (gfx950 HIP kernels behind the
reference's FastHigashi_Wrapper / parafac2_intergrative API)."""
__version__ = "0.1.0"
