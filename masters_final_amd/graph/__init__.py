"""Reference-compatible module namespace (MyML/MyML/graph)."""
