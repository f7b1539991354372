"""cyclosubfield-dbfv on MI355X: a gfx950 backend for the `hal::Backend` trait of
cyclosubfield-dbfv (RNS ring hot path: lazy radix-2/3 NTTs, Montgomery arithmetic, period<->CRT
transforms, base conversion, key-switch inner products), behind the C ABI of
include/csdbfv_hal.h. The package directory name is not a Python identifier; load it with
`importlib` (see `load_package` in __graft_entry__.py)."""
from .hal import I128, DeviceTensor, HalError, HipBackend, lib  # noqa: F401
from . import ring  # noqa: E402,F401
from . import bfv  # noqa: E402,F401
from . import dbfv  # noqa: E402,F401
from . import digitex  # noqa: E402,F401
from . import mfma_transform  # noqa: E402,F401
from . import homcrt  # noqa: E402,F401
from . import field  # noqa: E402,F401
from . import galoisring  # noqa: E402,F401
from . import keys  # noqa: E402,F401
from . import bootstrap  # noqa: E402,F401
