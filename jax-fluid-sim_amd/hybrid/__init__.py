"""MI355X-native hybrid DG hot path (host mirror of the reference hybrid package)."""
