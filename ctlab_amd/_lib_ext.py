"""ctypes declarations of the library's recon / learned-prox / batch entry points."""
import ctypes as C


def declare(L):
    pass
