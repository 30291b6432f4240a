"""Drop-in shim of the reference's ``InchingLiteInteger`` package entry points on the
eigenproblem path, backed by the MI355X implementation in ``inching_amd``."""
