"""ctf_amd — MI355X-native relativistic-MHD update of Mara (jzrake/ctf) behind Mara's
Lua-facing operator API.

* ``ctf_amd.native``     ctypes binding of the C ABI (``include/maragpu.h``, ``include/mara_host.h``)
* ``ctf_amd.mara``       the ``Mara`` verbs (start, set_fluid, ..., advance, get_timestep)
* ``ctf_amd.workloads``  the turbulence set-ups the measurements use
* ``ctf_amd.build``      hipcc build of ``libmaragpu.so``
"""
from . import native  # noqa: F401

__all__ = ["native"]
