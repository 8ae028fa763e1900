"""Mirror of FLUE.compiled (python/FLUE/compiled/__init__.py:1): the low-level ``flue_c``
module whose lower-case ``*_c`` functions wrap the C ABI of libflue_hip.so."""
from . import flue_c

__all__ = ["flue_c"]
