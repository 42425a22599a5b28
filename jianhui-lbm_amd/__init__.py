"""MI355X-native sparse D3Q19 MRT lattice-Boltzmann hot path (drop-in for jianhui-lbm).

The directory name carries a hyphen, so it is loaded by path (see ``load_package`` in
bench.py / tests/conftest.py); inside, modules import each other relatively.

* ``_capi``     -- ctypes binding of liblbm_mi355x.so (the C ABI in include/lbm_mi355x.h)
* ``solver``    -- ``SinglePhase`` / ``ColorGradient`` / ``Solute``: the reference's operations over
                   the C ABI
* ``serving``   -- ``PermeabilityServer``: continuous-batching permeability requests
* ``geometry``  -- porous-media generators and the reference's geometry file formats
* ``inputs``    -- the reference's positional INPUT.dat files
"""
from . import geometry, inputs  # noqa: F401
from ._capi import LBMError, build, declared_symbols, lib  # noqa: F401
from .serving import PermeabilityServer, request_stream  # noqa: F401
from .solver import ColorGradient, SinglePhase, Solute  # noqa: F401
