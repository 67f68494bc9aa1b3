"""MI355X (gfx950) HIP path of dxtb's GFN1/GFN2-xTB single point.

The calculators and parametrization records are imported on first use, so
that ``dxtb_amd._lib`` (the C-ABI library alone) loads without torch."""

_EXPORTS = {
    "Calculator": "calculators", "GFN1Calculator": "calculators", "GFN2Calculator": "calculators",
    "Param": "param", "GFN1_XTB": "param", "GFN2_XTB": "param",
}

__all__ = list(_EXPORTS)


def __getattr__(name):
    if name in _EXPORTS:
        import importlib

        return getattr(importlib.import_module(f".{_EXPORTS[name]}", __name__), name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
