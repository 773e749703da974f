"""gpuattract_2.0_amd — MI355X-native pose scoring for gpuATTRACT.

The directory name is not a Python identifier, so callers load it through
importlib.util.spec_from_file_location(...) under the module name
`gpuattract_amd` (tests/conftest.py, bench.py and __graft_entry__.py do so).
"""
from . import workload  # noqa: F401
