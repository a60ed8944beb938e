"""MI355X-native finite-PEPS contraction engine (libpeps_hip.so) and its host mirror."""
