"""Node trees on the host: metadata introspection through the C ABI, the
encoded-tree wire format (encoder), and node construction through setters.

Mirrors FastNoise2's Metadata / NodeData interface (include/FastNoise/
Metadata.h, src/FastNoise/Metadata.cpp of the reference):

* node ids are the registration order, member indices the metadata
  constructor order, both read from the library (``fnGetMetadata*``);
* ``NodeSpec`` is the NodeData analogue: a node name plus member values,
  where a member is a variable (float / int / enum), a node lookup or a
  hybrid (float constant or node);
* ``encode`` writes the reference's node stream (Metadata.cpp:161-308):
  ``u8 node id`` or ``0xFF u16 reference``, then only non-default members as
  ``MemberLookup{type: 3 LSB, index: 5 MSB}`` records (0 variable + i32,
  1 node lookups + children, 2 hybrid node + child, 3 hybrid float + f32,
  4 node end; consecutive ends merge into one record of index = run - 1),
  Base64-encoded with runs of 'A' written as '@' + Base64 digit of the run
  length minus 3 (Base64.h:29-59).

Member names match case-insensitively with spaces and underscores ignored;
per-axis members take the axis letter as a suffix (``amplitude_scaling_x``).
"""
import re
import struct

from . import _lib

_B64 = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"


def _norm(name):
    return re.sub(r"[^0-9a-z]", "", name.lower())


class MemberVariable:
    __slots__ = ("name", "index", "type", "dim", "default", "enum_names", "min", "max", "description")


class NodeMeta:
    """One registered node kind (Metadata of the reference)."""

    def __init__(self, lib, node_id):
        self.id = node_id
        self.name = lib.fnGetMetadataName(node_id).decode()
        self.description = lib.fnGetMetadataDescription(node_id).decode()
        self.groups = [lib.fnGetMetadataGroupName(node_id, g).decode()
                       for g in range(lib.fnGetMetadataGroupCount(node_id))]
        self.variables = []
        for v in range(lib.fnGetMetadataVariableCount(node_id)):
            m = MemberVariable()
            m.name = lib.fnGetMetadataVariableName(node_id, v).decode()
            m.index = v
            m.type = lib.fnGetMetadataVariableType(node_id, v)  # 0 float, 1 int, 2 enum
            m.dim = lib.fnGetMetadataVariableDimensionIdx(node_id, v)
            m.enum_names = [lib.fnGetMetadataEnumName(node_id, v, e).decode()
                            for e in range(lib.fnGetMetadataEnumCount(node_id, v))]
            if m.type == 0:
                m.default = lib.fnGetMetadataVariableDefaultFloat(node_id, v)
            else:
                m.default = lib.fnGetMetadataVariableDefaultIntEnum(node_id, v)
            m.min = lib.fnGetMetadataVariableMinFloat(node_id, v)
            m.max = lib.fnGetMetadataVariableMaxFloat(node_id, v)
            m.description = lib.fnGetMetadataVariableDescription(node_id, v).decode()
            self.variables.append(m)
        self.lookups = []
        for n in range(lib.fnGetMetadataNodeLookupCount(node_id)):
            self.lookups.append((lib.fnGetMetadataNodeLookupName(node_id, n).decode(),
                                 lib.fnGetMetadataNodeLookupDimensionIdx(node_id, n)))
        self.hybrids = []
        for h in range(lib.fnGetMetadataHybridCount(node_id)):
            self.hybrids.append((lib.fnGetMetadataHybridName(node_id, h).decode(),
                                 lib.fnGetMetadataHybridDimensionIdx(node_id, h),
                                 lib.fnGetMetadataHybridDefault(node_id, h)))
        self.keys = {}
        for m in self.variables:
            self.keys[self._key(m.name, m.dim)] = ("var", m.index)
        for idx, (name, dim) in enumerate(self.lookups):
            self.keys[self._key(name, dim)] = ("lookup", idx)
        for idx, (name, dim, _) in enumerate(self.hybrids):
            self.keys[self._key(name, dim)] = ("hybrid", idx)

    @staticmethod
    def _key(name, dim):
        return _norm(name) + ("xyzw"[dim] if dim >= 0 else "")

    def member(self, key):
        k = _norm(key)
        if k not in self.keys:
            raise KeyError("node %s has no member %r (members: %s)" % (self.name, key, sorted(self.keys)))
        return self.keys[k]


_registry = None


def registry():
    """All node kinds, indexed by id and by normalised name."""
    global _registry
    if _registry is None:
        lib = _lib.load()
        metas = [NodeMeta(lib, i) for i in range(lib.fnGetMetadataCount())]
        by_name = {_norm(m.name): m for m in metas}
        _registry = (metas, by_name)
    return _registry


def meta(name_or_id):
    metas, by_name = registry()
    if isinstance(name_or_id, int):
        return metas[name_or_id]
    k = _norm(name_or_id)
    if k not in by_name:
        raise KeyError("unknown node %r" % name_or_id)
    return by_name[k]


class NodeSpec:
    """A node and its member values (NodeData of the reference).

    ``NodeSpec("FractalFBm", source=NodeSpec("Simplex", feature_scale=200), octaves=5)``
    The same NodeSpec object used twice in a tree is encoded once and then
    referenced, as the reference does for shared nodes.
    """

    def __init__(self, name, **members):
        self.meta = meta(name)
        self.vars = {}
        self.lookups = {}
        self.hybrids = {}
        for k, v in members.items():
            self.set(k, v)

    def set(self, key, value):
        kind, idx = self.meta.member(key)
        if kind == "var":
            m = self.meta.variables[idx]
            if m.type == 2 and isinstance(value, str):
                names = [_norm(e) for e in m.enum_names]
                value = names.index(_norm(value))
            self.vars[idx] = value
        elif kind == "lookup":
            if not isinstance(value, NodeSpec):
                raise TypeError("node lookup %r needs a NodeSpec" % key)
            self.lookups[idx] = value
        else:
            self.hybrids[idx] = value
        return self

    # --- wire format -------------------------------------------------------
    def _var_bits(self, m, value):
        if m.type == 0:
            return struct.unpack("<i", struct.pack("<f", float(value)))[0]
        return int(value)

    def _encode(self, out, refs):
        if id(self) in refs:
            out.node_ref(refs[id(self)][0])
            return
        md = self.meta
        out.byte(md.id)
        for m in md.variables:
            if m.index in self.vars:
                bits = self._var_bits(m, self.vars[m.index])
                if bits != self._var_bits(m, m.default):
                    out.member(0, m.index)
                    out.i32(bits)
        if md.lookups:
            out.member(1, len(md.lookups))
            for idx in range(len(md.lookups)):
                child = self.lookups.get(idx)
                if child is None:
                    raise ValueError("node %s: lookup %r is not set" % (md.name, md.lookups[idx][0]))
                child._encode(out, refs)
        for idx, (_, _, default) in enumerate(md.hybrids):
            v = self.hybrids.get(idx)
            if isinstance(v, NodeSpec):
                out.member(2, idx)
                v._encode(out, refs)
            elif v is not None:
                if struct.pack("<f", float(v)) != struct.pack("<f", default):
                    out.member(3, idx)
                    out.raw(struct.pack("<f", float(v)))
        out.end()
        refs[id(self)] = (len(refs), self)

    def encode(self):
        """Encoded node tree string (SerialiseNodeData of the reference)."""
        out = _Stream()
        self._encode(out, {})
        out.flush()
        return b64encode(bytes(out.buf))

    # --- construction through the C ABI setters ------------------------------
    def create(self):
        """Builds the tree with fnNewFromMetadata + fnSet* (no wire format)."""
        from . import Node
        lib = _lib.load()
        made = {}

        def build(spec):
            if id(spec) in made:
                return made[id(spec)]
            n = Node._adopt(lib.fnNewFromMetadata(spec.meta.id, 0))
            for idx, v in spec.vars.items():
                m = spec.meta.variables[idx]
                ok = lib.fnSetVariableFloat(n.handle, idx, float(v)) if m.type == 0 else \
                    lib.fnSetVariableIntEnum(n.handle, idx, int(v))
                if not ok:
                    raise ValueError("cannot set %s.%s" % (spec.meta.name, m.name))
            for idx, child in spec.lookups.items():
                c = build(child)
                if not lib.fnSetNodeLookup(n.handle, idx, c.handle):
                    raise ValueError("cannot set lookup %d of %s" % (idx, spec.meta.name))
            for idx, v in spec.hybrids.items():
                if isinstance(v, NodeSpec):
                    ok = lib.fnSetHybridNodeLookup(n.handle, idx, build(v).handle)
                else:
                    ok = lib.fnSetHybridFloat(n.handle, idx, float(v))
                if not ok:
                    raise ValueError("cannot set hybrid %d of %s" % (idx, spec.meta.name))
            made[id(spec)] = n
            return n

        return build(self)


class _Stream:
    def __init__(self):
        self.buf = bytearray()
        self.pending_ends = 0

    def flush(self):
        if self.pending_ends:
            self.buf.append(4 | ((self.pending_ends - 1) << 3))
            self.pending_ends = 0

    def byte(self, v):
        self.flush()
        self.buf.append(v & 0xFF)

    def member(self, typ, index):
        self.byte(typ | (index << 3))

    def i32(self, v):
        self.flush()
        self.buf += struct.pack("<i", v)

    def raw(self, data):
        self.flush()
        self.buf += data

    def node_ref(self, ref):
        self.flush()
        self.buf.append(0xFF)
        self.buf += struct.pack("<H", ref)

    def end(self):
        self.pending_ends += 1
        if self.pending_ends >= 32:
            self.flush()


def b64encode(data):
    """Base64 with '@' runs of 3..66 'A' (Base64.h:29-59 of the reference)."""
    plain = []
    for k in range(0, len(data) - len(data) % 3, 3):
        t = (data[k] << 16) | (data[k + 1] << 8) | data[k + 2]
        plain += [_B64[(t >> s) & 63] for s in (18, 12, 6, 0)]
    rem = len(data) % 3
    if rem:
        t = data[-rem] << 16
        if rem == 2:
            t |= data[-1] << 8
        plain += [_B64[(t >> 18) & 63], _B64[(t >> 12) & 63],
                  _B64[(t >> 6) & 63] if rem == 2 else "=", "="]
    out = []
    k = 0
    while k < len(plain):
        run = 0
        while k + run < len(plain) and plain[k + run] == "A" and run < 66:
            run += 1
        if run >= 3:
            out.append("@" + _B64[run - 3])
            k += run
        elif run:
            out.append("A" * run)
            k += run
        else:
            out.append(plain[k])
            k += 1
    return "".join(out)


def b64decode(text):
    """Inverse of b64encode."""
    plain = []
    k = 0
    while k < len(text):
        if text[k] == "@":
            plain.append("A" * (_B64.index(text[k + 1]) + 3))
            k += 2
        else:
            plain.append(text[k])
            k += 1
    p = "".join(plain)
    out = bytearray()
    for k in range(0, len(p), 4):
        quad = p[k:k + 4]
        vals = [_B64.index(c) if c != "=" else 0 for c in quad]
        t = (vals[0] << 18) | (vals[1] << 12) | (vals[2] << 6) | vals[3]
        n = 3 - quad.count("=")
        out += bytes([(t >> 16) & 255, (t >> 8) & 255, t & 255][:n])
    return bytes(out)


def decode(encoded):
    """Parses an encoded node tree into NodeSpecs (DeserialiseNodeData of the
    reference, Metadata.cpp:311-353); shared nodes come back as the same
    NodeSpec object. Raises ValueError on malformed input."""
    data = b64decode(encoded)
    pos = [0]
    refs = []
    pending_ends = [0]

    def take(n):
        if pending_ends[0] or pos[0] + n > len(data):
            raise ValueError("truncated node stream")
        v = data[pos[0]:pos[0] + n]
        pos[0] += n
        return v

    def member():
        if pending_ends[0]:
            pending_ends[0] -= 1
            return 4, 0
        b = take(1)[0]
        typ, idx = b & 7, b >> 3
        if typ == 4:
            pending_ends[0] = idx
        return typ, idx

    metas, _ = registry()

    def node():
        nid = take(1)[0]
        if nid == 0xFF:
            ref = struct.unpack("<H", take(2))[0]
            if ref >= len(refs):
                raise ValueError("bad node reference")
            return refs[ref]
        if nid >= len(metas):
            raise ValueError("unknown node id %d" % nid)
        spec = NodeSpec(metas[nid].name)
        md = spec.meta
        typ, idx = member()
        while typ == 0:
            bits = struct.unpack("<i", take(4))[0]
            if idx < len(md.variables):
                m = md.variables[idx]
                spec.vars[idx] = struct.unpack("<f", struct.pack("<i", bits))[0] if m.type == 0 else bits
            typ, idx = member()
        if typ == 1:
            for k in range(idx):
                child = node()
                if k < len(md.lookups):
                    spec.lookups[k] = child
            typ, idx = member()
        while typ != 4:
            if typ == 2:
                child = node()
                if idx < len(md.hybrids):
                    spec.hybrids[idx] = child
            elif typ == 3:
                v = struct.unpack("<f", take(4))[0]
                if idx < len(md.hybrids):
                    spec.hybrids[idx] = v
            else:
                raise ValueError("bad member record")
            typ, idx = member()
        refs.append(spec)
        return spec

    return node()
