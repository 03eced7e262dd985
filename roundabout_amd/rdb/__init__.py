"""Python driver of the column-physics library (mirror of the reference's
``python/rdb`` package for the vertical-mixing path)."""

from ._ffi import (CLOSURE_PLAIN, CLOSURE_SURROGATE, GUARD_RANGE, GUARD_STRICT, GUARD_SURFACE, PREC_FP64,
                   PREC_SPLIT, Config, LibraryNotFoundError, MlpConfig, RdbError, default_config, load)
from ._model import ColumnEnsemble, ColumnPhysics
from .closures import KPP, PacanowskiPhilander, VmixAssembly

__all__ = ["ColumnPhysics", "ColumnEnsemble", "Config", "default_config", "load", "RdbError", "LibraryNotFoundError",
           "PacanowskiPhilander", "KPP", "VmixAssembly", "PREC_FP64", "PREC_SPLIT", "CLOSURE_PLAIN",
           "CLOSURE_SURROGATE", "GUARD_STRICT", "GUARD_SURFACE", "GUARD_RANGE", "MlpConfig"]
