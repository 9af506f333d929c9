"""Device buffers for the tests of the *_device entry points: the package's
own (libigl_amd.devmem), which go through the HIP runtime libgwn.so links."""
from libigl_amd.devmem import D2H, H2D, DeviceArray, Stream, hip, synchronize  # noqa: F401
