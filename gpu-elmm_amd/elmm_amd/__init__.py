"""MI355X-native ELMM LES time step: Python host mirror of the C ABI."""
from . import capi, cases  # noqa: F401
