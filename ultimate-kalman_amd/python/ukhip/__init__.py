"""Python host side of libukhip: ctypes bindings and seeded problem generators."""
from . import capi, problems  # noqa: F401
