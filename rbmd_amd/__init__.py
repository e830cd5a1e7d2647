"""MI355X-native RBMD force path (host side).

The compute path is the HIP library built from rbmd_amd/csrc into
rbmd_amd/_build/librbmd_mi355x.so and reached through its C ABI
(include/rbmd_mi355x.h).  This package holds the Python mirror of the
reference's seams (locator / force function / executioner), the input
readers and the synthetic-system generators.
"""
from .abi import Params, SysDesc, Thermo  # noqa: F401
from .config import RunConfig, load_json, run_config_from_json, salt_config, water_config  # noqa: F401
from .system import System, make_salt, make_water, read_data, rocksalt_lattice, special_lists, write_data  # noqa: F401

__version__ = "0.1.0"
