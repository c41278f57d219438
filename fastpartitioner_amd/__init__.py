"""FastPartitioner refinement and metric engine for MI355X (gfx950)."""
from .hypergraph import Hypergraph, colnet, workload, transpose, read_patoh, write_patoh  # noqa: F401
