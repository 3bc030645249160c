"""Write tests/golden/ref_vectors.npz from the reference's own functions.

Runs the reference's functions compiled into oracle/_ref/libkdn_ref.so
(`make -C oracle ref`, from its sobol.cpp, random.cpp, colour.cpp,
spectrum.cpp, energy.cpp, core/stopping.cpp and shared/image.cpp) on seeded
inputs and keeps inputs and outputs, so that the oracle's pins hold where the
reference tree is absent. Inputs come from a fixed seed: rerunning writes the
same file.

    python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libkdn_ref.so")
OUT = os.path.join(HERE, "ref_vectors.npz")

U64, D, F, I = C.c_uint64, C.c_double, C.c_float, C.c_int
P = C.c_void_p


def ref():
    L = C.CDLL(REF)
    L.kr_sobol_number.restype = D
    L.kr_sobol_number.argtypes = [U64, U64, U64]
    L.kr_hashed_number.restype = D
    L.kr_hashed_number.argtypes = [U64, U64, U64]
    L.kr_sampled_wavelengths.argtypes = [D, P, P]
    L.kr_colour_of.argtypes = [P, P, P, P]
    L.kr_spectrum_of.argtypes = [P, P, P, P]
    L.kr_escape_of.argtypes = [D, P, P]
    L.kr_effort_for.argtypes = [D, P, P]
    L.kr_split_ceiling.restype = I
    L.kr_split_ceiling.argtypes = [I]
    L.kr_fold_samples.argtypes = [P, I, I, P]
    L.kr_split_spread.argtypes = [P, I, I, P]
    L.kr_steps_between.restype = D
    L.kr_steps_between.argtypes = [D, D, F, I]
    L.kr_encoded_byte.restype = C.c_uint
    L.kr_encoded_byte.argtypes = [D, F, I]
    L.kr_write_ppm.argtypes = [I, I, P, F, I, C.c_char_p]
    L.kr_write_pfm.argtypes = [I, I, P, C.c_char_p]
    return L


def ptr(a):
    return a.ctypes.data_as(P)


class Lens(C.Structure):
    _fields_ = [("eye", C.c_double * 3), ("forward", C.c_double * 3), ("right", C.c_double * 3), ("up", C.c_double * 3),
                ("reach", C.c_double), ("aspect", C.c_double), ("width", C.c_int32), ("height", C.c_int32)]


def lens_struct(lens, width, height):
    L = Lens()
    L.eye[:], L.forward[:], L.right[:], L.up[:] = lens[0], lens[1], lens[2], lens[3]
    L.reach, L.aspect, L.width, L.height = lens[4], lens[5], width, height
    return L


def temporal_vectors(L):
    sys.path.insert(0, ROOT)
    from oracle import temporal_ref as T
    from tests import temporal_frames as TF
    L.kr_temporal_frame.argtypes = [I, I, P, P, I, I, I, P, P, P, I, I, I, P, P, P]
    frames = TF.sequence(seed=3, width=24, height=16, frames=8)
    out = {}
    bw = bh = bd = 0
    before = np.zeros((1, 1), T.PIXEL_DTYPE)
    for k, fr in enumerate(frames):
        W, H = fr["width"], fr["height"]
        px = np.zeros((H, W), T.PIXEL_DTYPE)
        img = np.zeros((H, W, 3), np.float32)
        L.kr_temporal_frame(W, H, C.byref(lens_struct(fr["lens"], W, H)), C.byref(lens_struct(fr["before_lens"], W, H)), int(fr["moved"]),
                            int(fr["restarted"]), int(fr["recounted"]), ptr(fr["parts"]), ptr(fr["record"]), ptr(fr["gbuffer"]), bw,
                            bh, bd, ptr(before), ptr(px), ptr(img))
        for name in ("parts", "record", "gbuffer"):
            out[f"temporal{k}_{name}"] = fr[name]
        out[f"temporal{k}_lens"] = np.array([*fr["lens"][0], *fr["lens"][1], *fr["lens"][2], *fr["lens"][3], fr["lens"][4],
                                             fr["lens"][5]])
        out[f"temporal{k}_flags"] = np.array([fr["moved"], fr["restarted"], fr["recounted"]], np.int32)
        out[f"temporal{k}_pixels"] = px
        out[f"temporal{k}_image"] = img
        before, bw, bh, bd = px, W, H, 1
    return out


def main():
    L = ref()
    rng = np.random.default_rng(20261204)
    out = {}
    # Sobol numbers: random streams, sample indices and every coordinate
    n = 4096
    streams = rng.integers(0, 2 ** 63, n, dtype=np.int64).astype(np.uint64) * np.uint64(2) + rng.integers(0, 2, n).astype(np.uint64)
    index = rng.integers(0, 1 << 20, n).astype(np.uint64)
    coord = rng.integers(0, 28, n).astype(np.uint64)
    out["sobol_stream"], out["sobol_index"], out["sobol_coordinate"] = streams, index, coord
    out["sobol_value"] = np.array([L.kr_sobol_number(int(s), int(i), int(c)) for s, i, c in zip(streams, index, coord)])
    # hashed numbers
    hc = rng.integers(0, 1 << 40, n).astype(np.uint64)
    out["hashed_stream"], out["hashed_index"], out["hashed_coordinate"] = streams, index, hc
    out["hashed_value"] = np.array([L.kr_hashed_number(int(s), int(i), int(c)) for s, i, c in zip(streams, index, hc)])
    # wavelengths at every cut and between them
    us = np.concatenate([np.arange(256) / 256.0, rng.random(1024)])
    nm = np.zeros((len(us), 4), np.float32)
    dens = np.zeros((len(us), 4), np.float32)
    for k, u in enumerate(us):
        L.kr_sampled_wavelengths(D(u), ptr(nm[k]), ptr(dens[k]))
    out["wave_u"], out["wave_nm"], out["wave_density"] = us, nm, dens
    # colourOf and spectrumOf at drawn wavelengths
    m = 512
    spec = rng.random((m, 4))
    rgbs = rng.random((m, 3)) * 2
    col = np.zeros((m, 3))
    back = np.zeros((m, 4))
    wn = nm[256:256 + m].astype(np.float64)
    wd = dens[256:256 + m].astype(np.float64)
    for k in range(m):
        a, b = np.ascontiguousarray(wn[k]), np.ascontiguousarray(wd[k])
        L.kr_colour_of(ptr(a), ptr(b), ptr(np.ascontiguousarray(spec[k])), ptr(col[k]))
        L.kr_spectrum_of(ptr(np.ascontiguousarray(rgbs[k])), ptr(a), ptr(b), ptr(back[k]))
    out["colour_nm"], out["colour_density"], out["colour_spectrum"], out["colour_rgb"] = wn, wd, spec, col
    out["spectrum_rgb"], out["spectrum_value"] = rgbs, back
    # GGX escape tables
    widths = np.array([1e-3, 0.01, 0.04, 0.09, 0.2, 0.35, 0.6, 1.0])
    tables = np.zeros((len(widths), 32))
    averages = np.zeros(len(widths))
    for k, w in enumerate(widths):
        avg = np.zeros(1)
        L.kr_escape_of(D(w), ptr(tables[k]), ptr(avg))
        averages[k] = avg[0]
    out["escape_width"], out["escape_table"], out["escape_average"] = widths, tables, averages
    # effortFor and splitCeiling
    tol = np.array([-1.0, 0.0, 0.05, 0.3, 0.5, 1.0, 1.5, 2.0, 3.0, 4.5, 6.0, 10.0, 25.0, 62.5, 80.0])
    eff = np.zeros((len(tol), 2), np.int32)
    for k, t in enumerate(tol):
        s, b = C.c_int(), C.c_int()
        L.kr_effort_for(D(t), C.byref(s), C.byref(b))
        eff[k] = (s.value, b.value)
    out["effort_tolerance"], out["effort"] = tol, eff
    out["ceiling_batch"] = np.arange(0, 130, dtype=np.int32)
    out["ceiling"] = np.array([L.kr_split_ceiling(int(b)) for b in out["ceiling_batch"]], np.int32)
    # foldedSample over samples alone (shared part = whole colour) and
    # splitSpread of the result: 64 pixels of 40 samples, batches of 8
    colours = np.exp(rng.normal(-1, 1.5, (64, 40, 3)))
    states = np.zeros((64, 18))
    spreads = np.zeros((64, 2), np.float32)
    for p in range(64):
        samples = np.concatenate([colours[p], colours[p]], axis=1).copy()
        L.kr_fold_samples(ptr(samples), 40, 8, ptr(states[p]))
        L.kr_split_spread(ptr(states[p]), 8, 1, ptr(spreads[p]))
    out["fold_colours"], out["fold_state"], out["fold_spread"] = colours, states, spreads
    # stepsBetween and encodedByte
    lows = np.exp(rng.normal(-1, 3, 2048))
    highs = lows + np.exp(rng.normal(-3, 3, 2048))
    exposure = rng.choice(np.array([1.0, 0.5, 2.3], np.float32), 2048)
    curve = rng.integers(0, 2, 2048).astype(np.int32)
    out["steps_low"], out["steps_high"], out["steps_exposure"], out["steps_curve"] = lows, highs, exposure, curve
    out["steps"] = np.array([L.kr_steps_between(a, b, F(e), int(c)) for a, b, e, c in zip(lows, highs, exposure, curve)])
    out["byte"] = np.array([L.kr_encoded_byte(a, F(e), int(c)) for a, e, c in zip(lows, exposure, curve)], np.uint8)
    # a PPM (filmic, exposure 1.5) and a PFM of a 7 x 5 image
    img = np.exp(rng.normal(-1, 1.5, (5, 7, 3))).astype(np.float32)
    tmp = os.path.join(HERE, "_tmp_image")
    L.kr_write_ppm(7, 5, ptr(img), F(1.5), 0, (tmp + ".ppm").encode())
    L.kr_write_pfm(7, 5, ptr(img), (tmp + ".pfm").encode())
    with open(tmp + ".ppm", "rb") as f:
        out["ppm"] = np.frombuffer(f.read(), np.uint8)
    with open(tmp + ".pfm", "rb") as f:
        out["pfm"] = np.frombuffer(f.read(), np.uint8)
    os.remove(tmp + ".ppm")
    os.remove(tmp + ".pfm")
    out["image"] = img
    # temporalFrame over a synthetic viewer sequence (tests/temporal_frames.py)
    out.update(temporal_vectors(L))
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    if not os.path.exists(REF):
        sys.exit(f"{REF} is not built: make -C oracle ref")
    main()
