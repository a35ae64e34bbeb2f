"""Minimal attribute-access config mapping.

The reference builds its runtime configuration with
``ml_collections.ConfigDict`` (``src/matcha/run.py:440-443``,
``src/matcha/example.py:358-361``).  That package is not a dependency here, so
this class gives the same attribute/item access over a plain dict and is
accepted anywhere a ConfigDict is; real ``ml_collections.ConfigDict`` objects
(or any mapping with ``__contains__``/``__getitem__``) are accepted too.
"""
from __future__ import annotations

from typing import Any, Mapping


class ConfigDict(dict):
    def __init__(self, data: Mapping | None = None, **kw):
        super().__init__()
        for k, v in dict(data or {}, **kw).items():
            self[k] = v

    def __setitem__(self, key, value):
        if isinstance(value, Mapping) and not isinstance(value, ConfigDict):
            value = ConfigDict(value)
        super().__setitem__(key, value)

    def __getattr__(self, key: str) -> Any:
        try:
            return self[key]
        except KeyError as exc:
            raise AttributeError(key) from exc

    def __setattr__(self, key: str, value: Any) -> None:
        self[key] = value

    def get(self, key, default=None):
        return super().get(key, default)


def cfg_get(cfg, key: str, default=None):
    """Read ``key`` from a mapping or attribute object."""
    if cfg is None:
        return default
    try:
        if key in cfg:
            return cfg[key]
        return default
    except TypeError:
        return getattr(cfg, key, default)
