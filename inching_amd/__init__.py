"""MI355X-native eigenproblem construction + diagonalisation for anisotropic network models.

Hot path: cell-list neighbours -> reverse Cuthill-McKee -> ANM 3x3 block Hessian ->
block SpMV / Chebyshev filter / Hotelling deflation -> TRLM / IRLM / JDM, all device work
in hand-written HIP kernels of libinching_hip.so (see include/inching.h).
"""
__version__ = "0.1.0"
