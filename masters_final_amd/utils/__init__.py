"""Reference-compatible module namespace (MyML/MyML/utils)."""
