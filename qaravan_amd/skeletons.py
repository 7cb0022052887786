"""Brickwork edge lists on 1D/2D lattices (src/qaravan/core/skeletons.py:9-64)."""


def index_from_coords(x: int, y: int, Lx: int) -> int:
    return y * Lx + x


def coords_from_index(idx: int, Lx: int):
    row, col = divmod(idx, Lx)
    return col, row


def horiz_filter(edge) -> bool:
    return abs(edge[0] - edge[1]) == 1


def even_filter(edge, vert=False, Lx=0) -> bool:
    if not vert:
        return edge[0] % 2 == 0
    _, row = coords_from_index(edge[0], Lx)
    return row % 2 == 0


def nn_edges(Lx: int, Ly: int = 1):
    """Nearest-neighbour edges of an Lx×Ly grid, row-major (skeletons.py:30-47)."""
    edges = []
    for row in range(Ly):
        for col in range(Lx):
            idx = index_from_coords(col, row, Lx)
            if col < Lx - 1:
                edges.append((idx, index_from_coords(col + 1, row, Lx)))
            if row < Ly - 1:
                edges.append((idx, index_from_coords(col, row + 1, Lx)))
    return edges


def brickwall_skeleton(Lx: int, Ly: int = 1, depth: int = 1):
    """Even horizontals, odd horizontals, even verticals, odd verticals, repeated
    ``depth`` times (skeletons.py:49-64)."""
    edges = nn_edges(Lx, Ly)
    h = [e for e in edges if horiz_filter(e)]
    v = [e for e in edges if not horiz_filter(e)]
    sk = [e for e in h if even_filter(e)] + [e for e in h if not even_filter(e)]
    sk += [e for e in v if even_filter(e, True, Lx)] + [e for e in v if not even_filter(e, True, Lx)]
    return sk * depth
