"""MI355X-native pressure-projection solver for LS-STAG (HIP kernels behind a C ABI).

The package directory name is not a Python identifier; load it with ``load_package()`` from
``__graft_entry__`` or ``importlib`` (``tests/conftest.py`` shows how).
"""
from . import lsx  # noqa: F401
from . import projection  # noqa: F401
