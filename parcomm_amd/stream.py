"""Model input into HBM: the reference CLI's dense BVA stream and safetensors
checkpoints, both written straight into a device-resident BVA.

- `read_description` / `load_bva_stream` take the converter's output: a JSON
  description (format "parcomm-dense-bva-stream-description", main.rs:116-128,
  146-153; written by modelconvert.py:1151-1191) and the raw stream of BVA
  entries in bucket order, which qf_stream_bva copies to the device through
  pinned double buffers.
- `load_safetensors` reads matrices out of safetensors files (an 8-byte
  little-endian header length, a JSON header of dtype / shape / data offsets,
  then the raw data) through memory maps and scatters each one into its
  buckets on the device (qf_encode_matrix, Layout::encode of layout.rs:115-135
  one matrix at a time). The tensors must already hold the entries:
  int8 for the F65521 build, or the raw 16-bit patterns of BF16 / int16 for the
  KoalaBear build (the converter's identity and BF16-bitcast conversions); the
  float-to-int8 quantisation of the converter is not part of the prover.
"""
import ctypes
import json
import os

import numpy as np

from . import DeviceBuffer, QfError, Shape

DESCRIPTION_FORMAT = "parcomm-dense-bva-stream-description"
STORAGE_DTYPES = {"i8": "int8", "i16": "int16"}


def read_description(path):
    """The converter's stream description: (shapes, storage dtype, the JSON)."""
    with open(path, "rb") as f:
        d = json.loads(f.read())
    if d.get("format") != DESCRIPTION_FORMAT or d.get("storage_dtype") not in STORAGE_DTYPES:
        raise ValueError("requires a parcomm-dense-bva-stream-description with storage_dtype i8 or i16")
    shapes = [Shape(int(r), int(c)) for r, c in d["shapes"]]
    return shapes, d["storage_dtype"], d


def check_description(params, description):
    """The stream must have been converted for this layout (main.rs:146-153):
    stream bytes = active entries x entry size, and the same logical length."""
    item = np.dtype(params.entry_dtype).itemsize
    if STORAGE_DTYPES[description["storage_dtype"]] != params.entry_dtype:
        raise ValueError("stream entry type does not match the field backend")
    logical = description.get("layout", {}).get("logical_entries")
    if description.get("stream_bytes") != item * params.active or logical != params.logical:
        raise ValueError("converter/ParComm layout mismatch")


def load_bva_stream(params, description, input_path, out=None):
    """The device BVA from a dense BVA stream (`input_path`, or "-" for stdin)."""
    check_description(params, description)
    item = np.dtype(params.entry_dtype).itemsize
    if out is None:
        out = DeviceBuffer(params.active * item, params.field)
    got = ctypes.c_uint64()
    status = params._L.qf_stream_bva(params._h, input_path.encode(), out.ptr, ctypes.byref(got))
    if status:
        raise QfError(status, params._L.qf_last_error().decode() + f" after {got.value} bytes")
    return out


SAFETENSORS_DTYPES = {"I8": np.int8, "U8": np.uint8, "I16": np.int16, "U16": np.uint16, "BF16": np.uint16,
                      "F16": np.uint16}


def safetensors_index(path):
    """name -> (numpy dtype, shape, absolute byte offset, byte length) of every tensor in a file."""
    with open(path, "rb") as f:
        n = int.from_bytes(f.read(8), "little")
        header = json.loads(f.read(n))
    out = {}
    for name, t in header.items():
        if name == "__metadata__":
            continue
        if t["dtype"] not in SAFETENSORS_DTYPES:
            raise ValueError(f"tensor {name}: dtype {t['dtype']} is not an 8- or 16-bit integer pattern")
        begin, end = t["data_offsets"]
        out[name] = (SAFETENSORS_DTYPES[t["dtype"]], tuple(t["shape"]), 8 + n + begin, end - begin)
    return out


def load_safetensors(params, files, names, out=None):
    """The device BVA of the named tensors, in the parameters' matrix order,
    from the safetensors `files`. A rank-1 tensor of length n is an n x 1
    matrix (modelconvert.py's rank1_matrix_mapping)."""
    if len(names) != len(params.shapes):
        raise ValueError("matrix count mismatch")
    index = {}
    for path in files:
        for name, rec in safetensors_index(path).items():
            index[name] = (path,) + rec
    item = np.dtype(params.entry_dtype).itemsize
    if out is None:
        out = DeviceBuffer(params.active * item, params.field)  # zeroed
    for i, (name, shape) in enumerate(zip(names, params.shapes)):
        path, dt, tshape, offset, nbytes = index[name]
        if np.dtype(dt).itemsize != item:
            raise ValueError(f"tensor {name}: {np.dtype(dt).itemsize}-byte entries, the backend takes {item}")
        rows, cols = (tshape[0], 1) if len(tshape) == 1 else (int(np.prod(tshape[:-1])), tshape[-1])
        if (rows, cols) != (shape.rows, shape.cols) or nbytes != rows * cols * item:
            raise ValueError(f"tensor {name}: shape {tshape} does not match matrix {i} ({shape.rows} x {shape.cols})")
        data = np.memmap(path, dtype=params.entry_dtype, mode="r", offset=offset, shape=(rows * cols,))
        params.encode_matrix(i, np.asarray(data), out)
        del data
    return out


def write_safetensors(path, tensors):
    """A safetensors file of numpy int8 / int16 / uint16 arrays (name -> array),
    for tests and small tools."""
    codes = {np.dtype(np.int8): "I8", np.dtype(np.uint8): "U8", np.dtype(np.int16): "I16", np.dtype(np.uint16): "U16"}
    header, blobs, off = {}, [], 0
    for name, a in tensors.items():
        a = np.ascontiguousarray(a)
        header[name] = {"dtype": codes[a.dtype], "shape": list(a.shape), "data_offsets": [off, off + a.nbytes]}
        blobs.append(a.tobytes())
        off += a.nbytes
    h = json.dumps(header).encode()
    h += b" " * ((8 - len(h) % 8) % 8)
    with open(path, "wb") as f:
        f.write(len(h).to_bytes(8, "little"))
        f.write(h)
        for b in blobs:
            f.write(b)
    return os.path.getsize(path)
