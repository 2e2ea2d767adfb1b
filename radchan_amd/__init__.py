"""MI355X radiative-transfer library: narrow-band correlated-k OERM Monte
Carlo for the radiative source term of a channel flow (a drop-in for
RadChan's src/montecarlo path on one GPU).

The product is the HIP library ``libradchan.so`` (C ABI in include/radchan.h);
this package binds it and provides parameter sets, synthetic tables and
fields.
"""
from . import lib
from .lib import VARIANT_BATCHED, VARIANT_PLAIN, VARIANT_PLAIN_BANDS, Params
from .params import base_params, config, reference_params
from .solver import Group, Solver, get_results, mc_gpu, set_dropin_modes, set_params, set_table_dir
from .tables import Tables, synthetic_nb

__all__ = [
    "lib", "Params", "VARIANT_PLAIN", "VARIANT_BATCHED", "VARIANT_PLAIN_BANDS", "base_params", "config", "reference_params",
    "Solver", "Group", "mc_gpu", "get_results", "set_params", "set_table_dir", "set_dropin_modes", "Tables", "synthetic_nb",
]
