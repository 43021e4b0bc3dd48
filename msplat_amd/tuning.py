"""Launch parameters of the rasterizer's blending kernels.

``cell``     edge, in 16x16 tiles, of the cells Gaussians are binned into;
``ppl_fwd``  pixels per lane of the blending forward (1, 2 or 4; 0 = default);
``ppl_bwd``  pixels per lane of the blending adjoint;
``cull``     skip, per tile, Gaussians that give none of its pixels alpha >= 1/255.

and, for a batch of cameras, how many views share one launch per stage.

Results do not depend on these (images, final_T and ncontrib are identical
for every choice; gradients differ only by float summation order).  The
defaults (cell 8, 2 pixels per lane forward, 4 in the adjoint) were the
fastest for C = 4 at 1920x1080 with 1M Gaussians on an MI355X; ``msplat_amd.autotune`` measures the
alternatives for a given image size and channel count and records the best
in the table below, which ``config_for`` consults.
"""

from __future__ import annotations

import dataclasses
import threading


@dataclasses.dataclass(frozen=True)
class BlendConfig:
    cell: int = 8
    ppl_fwd: int = 2
    ppl_bwd: int = 4
    cull: bool = True


DEFAULT = BlendConfig()

_lock = threading.Lock()
_table: dict[tuple[int, int, int], BlendConfig] = {}
_override: BlendConfig | None = None


def shape_key(W: int, H: int, C: int) -> tuple[int, int, int]:
    return int(W), int(H), int(C)


def config_for(W: int, H: int, C: int) -> BlendConfig:
    """The launch parameters to use for a W x H image with C channels."""
    with _lock:
        if _override is not None:
            return _override
        return _table.get(shape_key(W, H, C), DEFAULT)


def views_key(sizes, C: int) -> tuple:
    return ("views", tuple((int(w), int(h)) for w, h in sizes), int(C))


def config_for_views(sizes, C: int) -> BlendConfig:
    """Launch parameters for one many-view launch (a single view uses its (W, H, C) entry)."""
    if len(sizes) == 1:
        return config_for(sizes[0][0], sizes[0][1], C)
    with _lock:
        if _override is not None:
            return _override
        return _table.get(views_key(sizes, C), DEFAULT)


def record_views(sizes, C: int, cfg: BlendConfig) -> None:
    with _lock:
        _table[views_key(sizes, C)] = cfg


# Views per launch of a batch of cameras (msplat_amd.views): None renders the
# whole batch in one launch per stage; v splits it into consecutive groups of
# at most v views, one launch per stage each.
_per_launch: dict[tuple, int] = {}


def views_per_launch_for(sizes, C: int) -> int | None:
    with _lock:
        return _per_launch.get(views_key(sizes, C))


def record_views_per_launch(sizes, C: int, v: int | None) -> None:
    with _lock:
        if v is None:
            _per_launch.pop(views_key(sizes, C), None)
        else:
            _per_launch[views_key(sizes, C)] = int(v)


def record(W: int, H: int, C: int, cfg: BlendConfig) -> None:
    with _lock:
        _table[shape_key(W, H, C)] = cfg


def clear() -> None:
    with _lock:
        _table.clear()
        _per_launch.clear()


class override:
    """Context manager forcing one configuration (used by the autotuner and tests)."""

    def __init__(self, cfg: BlendConfig):
        self.cfg = cfg
        self.prev: BlendConfig | None = None

    def __enter__(self):
        global _override
        with _lock:
            self.prev, _override = _override, self.cfg
        return self.cfg

    def __exit__(self, *exc):
        global _override
        with _lock:
            _override = self.prev
        return False
