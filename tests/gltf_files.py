"""glTF 2.0 test files: a writer for synthetic .glb / .gltf documents and an
independent numpy reader used to check the C importer (meshLoadFromFile).

The reader follows the glTF 2.0 specification directly (GLB chunks, buffer
views with byteStride, sparse accessors, normalized integer components) and
shares no code with the library.
"""

import base64
import json
import struct

import numpy as np

COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
COMPONENT_ID = {np.dtype(v): k for k, v in COMPONENT.items()}
WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}
NORM_DIV = {np.int8: 127.0, np.uint8: 255.0, np.int16: 32767.0, np.uint16: 65535.0}


# ------------------------------------------------------------------ reading

def read(path):
    """Returns (document, list of buffer byte strings)."""
    raw = open(path, "rb").read()
    binary = None
    if raw[:4] == b"glTF":
        _, _, total = struct.unpack_from("<III", raw, 0)
        off, chunks = 12, []
        while off + 8 <= total:
            n, kind = struct.unpack_from("<II", raw, off)
            chunks.append((kind, raw[off + 8:off + 8 + n]))
            off += 8 + n
        doc = json.loads(chunks[0][1].decode().rstrip(" \0"))
        if len(chunks) > 1 and chunks[1][0] == 0x004E4942:
            binary = chunks[1][1]
    else:
        doc = json.loads(raw)
    buffers = []
    for b in doc.get("buffers", []):
        uri = b.get("uri")
        if uri is None:
            data = binary
        elif uri.startswith("data:"):
            data = base64.b64decode(uri.split(";base64,", 1)[1])
        else:
            import os
            import urllib.parse
            data = open(os.path.join(os.path.dirname(path), urllib.parse.unquote(uri)), "rb").read()
        buffers.append(data[:b["byteLength"]])
    return doc, buffers


def _view(doc, buffers, view_index, offset, dtype, width, count, strided=True):
    view = doc["bufferViews"][view_index]
    data = buffers[view["buffer"]]
    base = view.get("byteOffset", 0) + offset
    item = np.dtype(dtype).itemsize * width
    stride = view.get("byteStride", item) if strided else item
    out = np.empty((count, width), dtype)
    for i in range(count):
        out[i] = np.frombuffer(data, dtype, width, base + i * stride)
    return out


def accessor(doc, buffers, index, as_float=True):
    a = doc["accessors"][index]
    dtype = COMPONENT[a["componentType"]]
    width = WIDTH[a["type"]]
    count = a["count"]
    if "bufferView" in a:
        values = _view(doc, buffers, a["bufferView"], a.get("byteOffset", 0), dtype, width, count)
    else:
        values = np.zeros((count, width), dtype)
    if "sparse" in a:
        sp = a["sparse"]
        idx = _view(doc, buffers, sp["indices"]["bufferView"], sp["indices"].get("byteOffset", 0),
                    COMPONENT[sp["indices"]["componentType"]], 1, sp["count"], strided=False)[:, 0]
        val = _view(doc, buffers, sp["values"]["bufferView"], sp["values"].get("byteOffset", 0), dtype, width,
                    sp["count"], strided=False)
        values = values.copy()
        values[idx.astype(np.int64)] = val
    if not as_float:
        return values
    out = values.astype(np.float64)
    if a.get("normalized") and dtype in NORM_DIV:
        out = out / NORM_DIV[dtype]
    return out


def to_engine(v):
    """glTF +Y up to the engine's +Z up: (x, y, z) -> (x, -z, y)."""
    v = np.asarray(v, np.float64)
    return np.stack([v[:, 0], -v[:, 2], v[:, 1]], axis=1)


def node_matrix(node):
    """Local matrix (row-major math convention) of a glTF node."""
    if "matrix" in node:
        return np.array(node["matrix"], np.float64).reshape(4, 4).T
    t = node.get("translation", [0, 0, 0])
    x, y, z, w = node.get("rotation", [0, 0, 0, 1])
    s = node.get("scale", [1, 1, 1])
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r @ np.diag(s)
    m[:3, 3] = t
    return m


BASIS = np.array([[1, 0, 0, 0], [0, 0, -1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], np.float64)  # +90 degrees about X


def expected_primitive(doc, buffers, prim):
    """Engine-space positions, normals (None if absent) and indices after the
    winding alignment, computed from the file with float64 arithmetic."""
    attrs = prim["attributes"]
    pos = to_engine(accessor(doc, buffers, attrs["POSITION"]))
    nrm = to_engine(accessor(doc, buffers, attrs["NORMAL"])) if "NORMAL" in attrs else None
    if "indices" in prim:
        idx = accessor(doc, buffers, prim["indices"], as_float=False)[:, 0].astype(np.int64)
    else:
        idx = np.arange(len(pos))
    tri = idx.reshape(-1, 3).copy()
    if nrm is not None:
        face = np.cross(pos[tri[:, 1]] - pos[tri[:, 0]], pos[tri[:, 2]] - pos[tri[:, 0]])
        avg = nrm[tri].sum(axis=1)
        flip = ((face * avg).sum(1) < 0) & ((face * face).sum(1) > 1e-12) & ((avg * avg).sum(1) > 1e-12)
        tri[flip] = tri[flip][:, [0, 2, 1]]
    return pos, nrm, tri.ravel()


# ------------------------------------------------------------------ writing

class Builder:
    """Accumulates buffer views / accessors for one binary buffer."""

    def __init__(self):
        self.blob = bytearray()
        self.doc = {"asset": {"version": "2.0"}, "bufferViews": [], "accessors": []}

    def view(self, data: bytes, stride=None):
        while len(self.blob) % 4:
            self.blob.append(0)
        v = {"buffer": 0, "byteOffset": len(self.blob), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.blob.extend(data)
        self.doc["bufferViews"].append(v)
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, array, kind, normalized=False, view=None, offset=0, count=None, sparse=None):
        array = np.asarray(array)
        if view is None:
            view = self.view(array.tobytes())
        a = {"bufferView": view, "componentType": COMPONENT_ID[array.dtype], "count": count or len(array),
             "type": kind}
        if offset:
            a["byteOffset"] = offset
        if normalized:
            a["normalized"] = True
        if sparse is not None:
            a["sparse"] = sparse
        self.doc["accessors"].append(a)
        return len(self.doc["accessors"]) - 1

    def finish(self, path, container="glb", bin_name=None):
        doc = dict(self.doc)
        blob = bytes(self.blob)
        if container == "glb":
            doc["buffers"] = [{"byteLength": len(blob)}]
            js = json.dumps(doc).encode()
            js += b" " * (-len(js) % 4)
            blob += b"\0" * (-len(blob) % 4)
            body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
            open(path, "wb").write(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)
        elif container == "datauri":
            doc["buffers"] = [{"byteLength": len(blob),
                               "uri": "data:application/octet-stream;base64," + base64.b64encode(blob).decode()}]
            open(path, "w").write(json.dumps(doc))
        else:
            import os
            import urllib.parse
            doc["buffers"] = [{"byteLength": len(blob), "uri": urllib.parse.quote(bin_name)}]
            open(os.path.join(os.path.dirname(path), bin_name), "wb").write(blob)
            open(path, "w").write(json.dumps(doc))
