"""MI355X-native engine of the fast bidirectional vector wave propagation method (BWPM).

Drop-in for the slab-stepping path of the reference package `bwpm`: the same modules and types
(`fast`, `volume3d`, `fpwpm`, `bidirectional`, `interface`, `adapter`, `results`) over hand-written
gfx950 HIP kernels behind the C ABI of include/bwpm_hip.h (`_core`), plus batched and served solves
(`serve`). Conventions as the reference: lengths in mm, wavelengths in µm, fields ∝ exp(i(k·r − ωt)),
Fresnel Convention A, TE = s, TM = p.
"""

__version__ = "0.1.0"
