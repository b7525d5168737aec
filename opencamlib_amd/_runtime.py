"""One HIP runtime per process.

PyTorch ships its own copy of the HIP runtime (libamdhip64.so.7) and loads it
by path.  If the native libraries here were loaded first, the dynamic linker
would bind them to the system copy and PyTorch would then load a second
runtime into the process; two runtimes sharing one device fail.  Importing
torch first (when it is installed) makes the native libraries bind to the
copy already loaded, by soname.
"""


def preload_hip_runtime() -> None:
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
