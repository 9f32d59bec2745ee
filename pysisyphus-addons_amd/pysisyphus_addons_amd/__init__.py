"""MI355X-native density-fitting Coulomb path, drop-in for
pysisyphus_addons.diabatization.intor and pysisyphus_addons.intor.intor.

All integral and linear-algebra work runs in libdfhip.so (HIP, gfx950);
importing a submodule that needs it raises if the library or a GPU is
missing -- there is no CPU fallback.
"""
