"""MI355X (gfx950) drop-in for PowerSpectra.jl's mode-coupling and covariance
hot path.  The compute path is the HIP library libpsx_hip.so (C ABI in
include/powerspectra_hip.h); this package mirrors the reference's Julia API
on top of it (names as exported by the reference's src/PowerSpectra.jl:40-51).
"""
from . import _abi
from ._abi import (ArgumentError, AxisError, BudgetError, Context, DeviceError, DomainError,
                   PowerSpectraError, Server, SingularException, default_context)
from .alm import Alm, alm2cl
from .beam import BandedMatrix, ku, quickpolW, quickpolXi
from .blockspectralmatrix import (BlockSpectralMatrix, block_vector, getblock, hcat, hvcat, solve,
                                  spectra, vcat)
from .covariance import coupledcov, coupledcovTEEE, decouple_covmat
from .master import binning_matrix, maskedalm2spectra, master
from .modecoupling import mask, mcm, mcm_binned, mcm_from_spectrum, mcm_set, scale
from .sht import HealpixMap, alm2map, alm2map_pol, map2alm, map2alm_pol, nside2lmax, nside2npix
from .spectralarray import (SpectralArray, SpectralVector, inv, parent, spectralones,
                            spectralzeros)
from .workspace import (ConstantDict, CovarianceWorkspace, CovField, SpectralWorkspace,
                        SpectrumName)

__all__ = [
    "Alm", "ArgumentError", "AxisError", "BandedMatrix", "BlockSpectralMatrix", "BudgetError", "ConstantDict",
    "Context", "CovField", "CovarianceWorkspace", "DeviceError", "DomainError", "HealpixMap",
    "PowerSpectraError", "Server", "SingularException", "SpectralArray", "SpectralVector",
    "SpectralWorkspace", "SpectrumName", "alm2cl", "alm2map", "alm2map_pol", "binning_matrix",
    "block_vector", "coupledcov",
    "coupledcovTEEE", "decouple_covmat", "default_context", "getblock", "hcat", "hvcat", "inv",
    "map2alm", "map2alm_pol", "mask", "maskedalm2spectra", "master", "mcm", "mcm_binned",
    "mcm_from_spectrum",
    "mcm_set", "nside2lmax", "nside2npix",
    "ku", "parent", "quickpolW", "quickpolXi", "scale", "solve", "spectra", "spectralones", "spectralzeros", "vcat",
]
