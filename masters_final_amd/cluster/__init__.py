"""Reference-compatible module namespace (MyML/MyML/cluster)."""
