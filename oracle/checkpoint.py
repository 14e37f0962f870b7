"""CPU restatement of the reference's runtime-checkpoint file format (pure Python, struct module).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.  Follows
src/genetic_algorithm/device/slab_runtime.cu: the checksum ComputeRuntimeCheckpointChecksum (:1206-1248),
the identity hash of TryDownloadBoundaryCheckpointPayload (:1282-1296), the writer
TryWriteRuntimeCheckpointAtomically (:1934-2007) and the reader TryReadRuntimeCheckpoint (:2009-2081).
The configuration structs are written whole as their LP64 images; their sizes and offsets are pinned
against the reference headers compiled into oracle/_ref (tests/test_checkpoint_format.py).  The genome
records hold fp32 slots, which the genome layout version says ("FP32"; the reference's fp16 genomes are 1).
"""
import struct

MAGIC = 0x4B4350474147524E          # slab_runtime.cu:1184
SCHEMA_VERSION = 3                  # slab_runtime.hpp:86
GENOME_LAYOUT_VERSION = 0x32335046  # b"FP32": fp32 genome slots (the reference's fp16 layout is 1, slab_runtime.hpp:87)
REFERENCE_GENOME_LAYOUT_VERSION = 1
PHASE_PRE_RECOMBINATION = 1         # slab_runtime.hpp:89-91
RELEASE_CAPACITY = 4739             # training_data.hpp:22
FNV_PRIME = 1099511628211
FNV_OFFSET_BASIS = 14695981039346656037   # the published FNV-1a 64 offset basis
REFERENCE_FNV_START = 1469598103934665603  # what slab_runtime.cu:1207 and :1283 start from
MASK = (1 << 64) - 1

# LP64 struct images (sizes; GenerationAssemblyConfig has 4 padding bytes at 12, the injection 7 at 1)
WORD_COUNTS_BYTES, ASSEMBLY_BYTES, INJECTION_BYTES = 56, 40, 24
RUNTIME_CONFIG_BYTES, SLAB_LAYOUT_BYTES, GRID_SHAPE_BYTES = 48, 32, 24


def fnv1a64(data, start=REFERENCE_FNV_START):
    h = start
    for b in bytes(data):
        h = ((h ^ b) * FNV_PRIME) & MASK
    return h


def pack_word_counts(training, actions, initial, step, period, radius0, growth):
    return struct.pack("<7Q", training, actions, initial, step, period, radius0, growth)


def pack_assembly(radius, rank_exponent, t1, t2, t3, mp, ms, row_p, padding=b"\0" * 4):
    return struct.pack("<Qf", radius, rank_exponent) + padding + struct.pack("<6f", t1, t2, t3, mp, ms, row_p)


def unpack_assembly(b):
    radius, rank = struct.unpack_from("<Qf", b, 0)
    return (radius, rank) + struct.unpack_from("<6f", b, 16)


def pack_injection(enabled, first, count, padding=b"\0" * 7):
    return bytes([1 if enabled else 0]) + padding + struct.pack("<2Q", first, count)


def pack_grid_shape(rows, cols):
    return struct.pack("<3Q", rows, cols, rows * cols)


def identity_hash(word_counts, history=None, catalog=None):
    """word_counts: the 56-byte image; history: None or (generations, p99); catalog: None or [n][5] letters."""
    data = bytearray(word_counts)
    if history is not None:
        gens, p99 = history
        data += struct.pack("<Q", len(gens))
        for g, p in zip(gens, p99):
            data += struct.pack("<Qf", g, p)
    if catalog is not None:
        data += struct.pack("<Q", len(catalog))
        for w in catalog:
            data += bytes(int(x) for x in w)
    return fnv1a64(data)


def checksum(c):
    """c: dict with the fields of RuntimeCheckpoint (as read() returns)."""
    d = bytearray()
    d += struct.pack("<IIIQI", c["schema_version"], c["genome_layout_version"], c["resume_phase"],
                     c["identity_hash"], c["generation_seed"])
    d += c["word_counts"]
    d += struct.pack("<Q", len(c["release_generations"]))
    for g, p in zip(c["release_generations"], c["release_p99"]):
        d += struct.pack("<Qf", g, p)
    d += c["assembly"] + c["injection"] + c["runtime_config"] + c["slab_layout"] + b"".join(c["shapes"])
    d += struct.pack("<QQ", c["generation_index"], len(c["slots"]))
    for s, f, e, h in zip(c["slots"], c["fitness"], c["evals"], c["has"]):
        d += struct.pack("<IfIB", s, f, e, h)
    d += struct.pack("<Q", len(c["pairs"]))
    for a, b in c["pairs"]:
        d += struct.pack("<II", a, b)
    d += struct.pack("<Q", len(c["records"]))
    for org, genome in c["records"]:
        d += struct.pack("<IQ", org, len(genome)) + bytes(genome)
    return fnv1a64(d)


def write(c):
    """The file bytes of checkpoint dict c (its "checksum" field is written as given)."""
    d = bytearray(struct.pack("<QIIQQII", MAGIC, c["schema_version"], c["genome_layout_version"], c["checksum"],
                              c["identity_hash"], c["resume_phase"], c["generation_seed"]))
    d += c["word_counts"]
    d += struct.pack("<Q", len(c["release_generations"]))
    for g, p in zip(c["release_generations"], c["release_p99"]):
        d += struct.pack("<Qf", g, p)
    d += c["assembly"] + c["injection"] + c["runtime_config"] + c["slab_layout"] + b"".join(c["shapes"])
    d += struct.pack("<QQ", c["generation_index"], len(c["slots"]))
    for s, f, e, h in zip(c["slots"], c["fitness"], c["evals"], c["has"]):
        d += struct.pack("<IfIB", s, f, e, h)
    d += struct.pack("<Q", len(c["pairs"]))
    for a, b in c["pairs"]:
        d += struct.pack("<II", a, b)
    d += struct.pack("<Q", len(c["records"]))
    for org, genome in c["records"]:
        d += struct.pack("<IQ", org, len(genome)) + bytes(genome)
    return bytes(d)


class Invalid(Exception):
    pass


def read(data):
    """Parses file bytes as TryReadRuntimeCheckpoint does; raises Invalid where it returns false."""
    pos = 0

    def take(fmt):
        nonlocal pos
        size = struct.calcsize(fmt)
        if pos + size > len(data):
            raise Invalid("truncated")
        v = struct.unpack_from(fmt, data, pos)
        pos += size
        return v

    def raw(n):
        nonlocal pos
        if pos + n > len(data):
            raise Invalid("truncated")
        v = bytes(data[pos:pos + n])
        pos += n
        return v

    magic, schema, layout, chk, ident, phase, seed = take("<QIIQQII")
    if magic != MAGIC:
        raise Invalid("magic")
    c = {"schema_version": schema, "genome_layout_version": layout, "checksum": chk, "identity_hash": ident,
         "resume_phase": phase, "generation_seed": seed, "word_counts": raw(WORD_COUNTS_BYTES)}
    (releases,) = take("<Q")
    if releases > RELEASE_CAPACITY:
        raise Invalid("release history")
    rel = [take("<Qf") for _ in range(releases)]
    c["release_generations"] = [g for g, _ in rel]
    c["release_p99"] = [p for _, p in rel]
    c["assembly"] = raw(ASSEMBLY_BYTES)
    c["injection"] = raw(INJECTION_BYTES)
    c["runtime_config"] = raw(RUNTIME_CONFIG_BYTES)
    c["slab_layout"] = raw(SLAB_LAYOUT_BYTES)
    c["shapes"] = [raw(GRID_SHAPE_BYTES) for _ in range(3)]
    c["generation_index"], count = take("<QQ")
    if count == 0:
        raise Invalid("empty generation")
    rows = [take("<IfIB") for _ in range(count)]
    c["slots"] = [r[0] for r in rows]
    c["fitness"] = [r[1] for r in rows]
    c["evals"] = [r[2] for r in rows]
    c["has"] = [r[3] for r in rows]
    (children,) = take("<Q")
    if children == 0:
        raise Invalid("empty plan")
    c["pairs"] = [take("<II") for _ in range(children)]
    (live,) = take("<Q")
    if live > count:
        raise Invalid("live count")
    stride = struct.unpack_from("<Q", c["slab_layout"], 8)[0]
    c["records"] = []
    for _ in range(live):
        org, nbytes = take("<IQ")
        if nbytes != stride:
            raise Invalid("genome size")
        c["records"].append((org, raw(nbytes)))
    if schema != SCHEMA_VERSION or layout != GENOME_LAYOUT_VERSION or chk != checksum(c):
        raise Invalid("version or checksum")
    return c


def slab_layout(c):
    """(action_count, slot_stride_bytes, slot_count, slab_bytes)."""
    return struct.unpack("<4Q", c["slab_layout"])


def runtime_config(c):
    """(slab budget, generation budget, host spill budget, action_count, population ceiling, grid columns)."""
    return struct.unpack("<6Q", c["runtime_config"])


def grid_shapes(c):
    return [struct.unpack("<3Q", s) for s in c["shapes"]]
