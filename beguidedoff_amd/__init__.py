"""beguidedoff_amd — the extensive-mode off-target scan of BEguidedOFF on
AMD Instinct MI355X (gfx950).

Public surface mirrors the reference package (beguidedoff/__init__.py:19-54):
``beguidedoff(mode="extensive", ...)`` returns the GPU finder, whose
``search()`` takes the same inputs and returns the same 12-column table.
Importing the package loads neither torch nor the HIP library; the kernel
library is loaded when a finder is created.
"""
from .types import GuideInput, OffTarget

__all__ = ["beguidedoff", "beguidedofffast", "GuideInput", "OffTarget"]
__version__ = "0.1.0"


def beguidedoff(*args, mode: str = "extensive", **kwargs):
    """Create an off-target finder.

    ``mode="extensive"`` (default) returns :class:`beguidedoff_amd.finder.beguidedoff`;
    ``mode="fast"`` returns :class:`beguidedoff_amd.fast.beguidedofffast`, the
    reference's anchored search computed on the GPU without an FM-index.
    """
    mode = str(mode).lower()
    if mode == "extensive":
        from .finder import beguidedoff as _extensive

        return _extensive(*args, **kwargs)
    if mode == "fast":
        from .fast import beguidedofffast as _fast

        return _fast(*args, **kwargs)
    raise ValueError(f"Unknown mode {mode!r}; expected 'extensive' or 'fast'.")


def beguidedofffast(*args, **kwargs):
    """Shorthand for ``beguidedoff(mode="fast", ...)`` (beguidedoff/__init__.py:47-54)."""
    from .fast import beguidedofffast as _fast

    return _fast(*args, **kwargs)
