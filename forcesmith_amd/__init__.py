"""MI355X-native force-matching evaluator (placeholder; filled in below)."""
