"""MI355X-native sampling hot path: rank-width frontier contraction, renormalisation and CDF draw."""
