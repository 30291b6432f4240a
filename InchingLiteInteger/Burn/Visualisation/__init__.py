"""Drop-in module path of the reference package (see InchingLiteInteger/__init__.py)."""
