"""MI355X (gfx950) implementation of quasar's 2D constrained-transport ideal-MHD
step: hand-written HIP kernels behind the C ABI in include/quasar_mhd.h, with
``quasar_amd.mhd`` mirroring the reference's ``_core.mhd`` module."""

__all__ = ["mhd"]
