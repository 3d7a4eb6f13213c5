"""The messages of the reference's proto/uvlm.proto (lines 4-100) declared to
the Python protobuf runtime, without protoc: an independent decoder for the
host library's hand-written wire encoding.  Test infrastructure only.

Also reads the record streams the host library writes (a base-128 varint
byte length, then the message).
"""
from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

F = descriptor_pb2.FieldDescriptorProto
OPT, REQ, REP = F.LABEL_OPTIONAL, F.LABEL_REQUIRED, F.LABEL_REPEATED
DBL, U64, U32, MSG = F.TYPE_DOUBLE, F.TYPE_UINT64, F.TYPE_UINT32, F.TYPE_MESSAGE

# (message, [(field, number, label, type, message type)])
MESSAGES = [
    ("Point", [("x", 1, OPT, DBL, None), ("y", 2, OPT, DBL, None), ("z", 3, OPT, DBL, None)]),
    ("Wing", [("cols", 1, OPT, U64, None), ("rows", 2, OPT, U64, None),
              ("points", 3, REP, MSG, "Point"), ("origin", 4, OPT, MSG, "Point"),
              ("chord", 5, OPT, DBL, None), ("span", 6, OPT, DBL, None)]),
    ("VortexRing", [("gamma", 1, OPT, DBL, None), ("nodes", 2, REP, MSG, "Point"),
                    ("nodes0", 3, REP, MSG, "Point")]),
    ("MorphingVelocity", [("center", 1, OPT, MSG, "Point"), ("nodes", 2, REP, MSG, "Point"),
                          ("freestream", 3, OPT, MSG, "Point")]),
    ("Snapshot2", [("t", 1, OPT, DBL, None),
                   ("container_shapes", 2, REP, MSG, "Snapshot2.ContainerShape"),
                   ("vortices", 3, REP, MSG, "VortexRing"),
                   ("morphing_velocities", 4, REP, MSG, "MorphingVelocity")]),
    ("VortexSheet", [("t", 1, OPT, DBL, None), ("num", 2, REQ, U32, None),
                     ("rows", 3, REQ, U32, None), ("cols", 4, REQ, U32, None),
                     ("vortices", 5, REP, MSG, "VortexRing")]),
    ("AllVortexSheets", [("t", 1, OPT, DBL, None), ("wing", 2, OPT, MSG, "VortexSheet"),
                         ("wake", 3, OPT, MSG, "VortexSheet")]),
]
NESTED = {"Snapshot2": ("ContainerShape", [
    ("rows", 1, REQ, U32, None), ("cols", 2, REQ, U32, None), ("id", 3, OPT, U32, None),
    ("origin", 4, OPT, MSG, "Point"), ("chord", 5, OPT, DBL, None),
    ("span", 6, OPT, DBL, None)])}


def _add_fields(msg, fields):
    for name, num, label, typ, ref in fields:
        f = msg.field.add(name=name, number=num, label=label, type=typ)
        if ref:
            f.type_name = ".UVLM.proto." + ref


def _build():
    fd = descriptor_pb2.FileDescriptorProto(name="uvlm_test.proto", package="UVLM.proto",
                                            syntax="proto2")
    for name, fields in MESSAGES:
        m = fd.message_type.add(name=name)
        _add_fields(m, fields)
        if name in NESTED:
            nname, nfields = NESTED[name]
            _add_fields(m.nested_type.add(name=nname), nfields)
    pool = descriptor_pool.DescriptorPool()
    pool.Add(fd)
    get = message_factory.GetMessageClass
    return {name: get(pool.FindMessageTypeByName("UVLM.proto." + name)) for name, _ in MESSAGES}


CLASSES = _build()


def parse(name, data: bytes):
    m = CLASSES[name]()
    m.ParseFromString(data)
    return m


def read_records(path):
    """The messages of a record stream, as bytes."""
    data = open(path, "rb").read()
    out, k = [], 0
    while k < len(data):
        n, shift = 0, 0
        while True:
            b = data[k]
            k += 1
            n |= (b & 0x7F) << shift
            shift += 7
            if not b & 0x80:
                break
        out.append(data[k:k + n])
        k += n
    return out
