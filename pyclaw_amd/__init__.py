"""pyclaw_amd -- MI355X-native time-stepping hot path with PyClaw's Python API."""
