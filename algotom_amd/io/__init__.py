"""File output used by the reconstruction path (algotom/io/loadersaver.py)."""
