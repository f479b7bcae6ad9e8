"""Reference-compatible module namespace (MyML/MyML/EAC)."""
