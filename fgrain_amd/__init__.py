"""MI355X-native (gfx950, HIP) photon-transport + transmittance-scan engine with
fgrain's library API (GuyNguyen/fgrain src/lib.rs).  The compute lives in
libfgrain_hip.so behind the C ABI of include/fgrain_hip.h."""
from .api import *  # noqa: F401,F403
from .api import __all__  # noqa: F401
