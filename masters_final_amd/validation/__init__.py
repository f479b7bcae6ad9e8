"""Reference-compatible module namespace (MyML/MyML/validation): clustering
accuracy indices over a contingency table counted on the GPU."""
