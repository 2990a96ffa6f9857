"""pyglu (work in progress)."""
