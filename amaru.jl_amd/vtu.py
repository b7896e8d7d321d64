"""VTU output of a model and its nodal fields: ``save_vtu`` follows the
uncompressed form of the reference's ``save_vtu`` (src/mesh/io.jl:167-276,
``get_array_node!`` :126-163): an UnstructuredGrid piece with the node
coordinates, the cells (0-based connectivity in the reference's node order,
offsets, VTK cell types), then point data and cell data arrays, each an
ASCII DataArray (floats written as Float32 with ``%20.10e``, integers
followed by two spaces, one row per point or cell).

The reference's HEX20 and TET10 node orders are VTK's quadratic hexahedron
and tetrahedron orders, so the connectivity is written as stored.
"""
import numpy as np

VTK_TYPE = {"HEX8": 12, "HEX20": 25, "TET10": 24}


def _array(name, A):
    A = np.asarray(A)
    if A.ndim == 1:
        A = A[:, None]
    isfloat = np.issubdtype(A.dtype, np.floating)
    dtype = "Float64" if isfloat else {np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64"}.get(A.dtype, "Int64")
    rows = []
    for row in A:
        if isfloat:
            rows.append("".join("%20.10e" % float(np.float32(v)) for v in row))
        else:
            rows.append("".join(f"{int(v)}  " for v in row))
    head = f'<DataArray type="{dtype}" Name="{name}" NumberOfComponents="{A.shape[1]}" format="ascii">'
    return head + "\n" + "\n".join(rows) + "\n</DataArray>"


def save_vtu(model, filename, node_data=None, elem_data=None, desc=""):
    """Writes the model's mesh with ``node_data`` (name -> nnodes or
    nnodes x k array, e.g. "U" and the recovered nodal fields) and
    ``elem_data`` (name -> ne or ne x k)."""
    mesh = model.mesh
    X = np.asarray(mesh.coords, dtype=float)
    conn = np.asarray(mesh.conn, dtype=np.int32)
    ne, nen = conn.shape
    parts = ['<?xml version="1.0" encoding="UTF-8"?>', f"<!-- {desc} -->",
             '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
             "<UnstructuredGrid>", f'<Piece NumberOfPoints="{len(X)}" NumberOfCells="{ne}">',
             "<Points>", _array("Points", X), "</Points>", "<Cells>",
             _array("connectivity", conn.reshape(-1).astype(np.int32)),
             _array("offsets", (np.arange(1, ne + 1) * nen).astype(np.int32)),
             _array("types", np.full(ne, VTK_TYPE[mesh.cellshape], dtype=np.int32)), "</Cells>"]
    for tag, data, n in (("PointData", node_data, len(X)), ("CellData", elem_data, ne)):
        if data:
            parts.append(f"<{tag}>")
            for name, A in data.items():
                A = np.asarray(A)
                if A.shape[0] != n:
                    raise ValueError(f"save_vtu: {name} has {A.shape[0]} rows, expected {n}")
                parts.append(_array(name, A))
            parts.append(f"</{tag}>")
    parts += ["</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    with open(filename, "w") as f:
        f.write("\n".join(parts) + "\n")
