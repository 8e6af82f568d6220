"""cadise_amd: MI355X (gfx950) wavefront path tracer, a drop-in for the
path-tracing renderer of xh5a5n6k6/cadise (renderer "sampling" + estimator
"pt" + accelerator "bvh"). The compute path is libmpt.so (include/mpt.h);
this package is its Python binding plus the seeded scene generators."""
from .native import MptError, load_library  # noqa: F401
