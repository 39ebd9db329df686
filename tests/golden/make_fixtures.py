"""Write the golden fixtures from the reference's own regression test data.

Run here only (it reads /root/reference/tests); the GPU box uses the files it
writes, which are committed:
  tests/golden/srcimg/<name>.pam   inputs: the reference's tests/srcimg PNGs
                                   converted the way the reference's PNG reader
                                   presents them to the library
                                   (src/imagew-png.c:266-401): palette ->
                                   RGB(A), tRNS -> alpha, other gray depths kept
                                   via MAXVAL (2^depth-1)
  tests/golden/expected.npz        the reference's expected outputs
                                   (tests/expected/*.png) decoded to canonical
                                   RGBA at their sample depth, with colour
                                   zeroed where alpha is 0
  tests/golden/cases.json          command-line arguments per case, transcribed
                                   from the reference's tests/runtest

None of the source PNGs carries gAMA, so PAM's default sRGB input colour space
matches the PNGs' (sRGB chunk or none).  Cases whose inputs carry bKGD or sBIT,
or that use codecs other than PNG, are left out (PAM cannot carry them).
"""
import json
import os
import struct
import sys
import zlib

import numpy as np

REF = "/root/reference/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_png(path):
    """Minimal PNG decoder: returns dict(width, height, bit_depth, color_type,
    rows (list of numpy arrays of samples), plte, trns)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos = 8
    idat = b""
    info = {"plte": None, "trns": None}
    while pos < len(data):
        n, t = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if t == b"IHDR":
            w, h, bd, ct, _, _, il = struct.unpack(">IIBBBBB", body)
            info.update(width=w, height=h, bit_depth=bd, color_type=ct, interlace=il)
        elif t == b"PLTE":
            info["plte"] = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif t == b"tRNS":
            info["trns"] = body
        elif t == b"IDAT":
            idat += body
        elif t == b"IEND":
            break
    raw = zlib.decompress(idat)
    w, h, bd, ct = info["width"], info["height"], info["bit_depth"], info["color_type"]
    nch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ct]
    passes = [(0, 0, 1, 1)] if info["interlace"] == 0 else [
        (0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
    out = np.zeros((h, w, nch), np.uint32)
    p = 0
    bpp = max(1, nch * bd // 8)
    for (x0, y0, dx, dy) in passes:
        pw = (w - x0 + dx - 1) // dx
        ph = (h - y0 + dy - 1) // dy
        if pw == 0 or ph == 0:
            continue
        rowbytes = (pw * nch * bd + 7) // 8
        prev = bytearray(rowbytes)
        for y in range(ph):
            f = raw[p]
            line = bytearray(raw[p + 1:p + 1 + rowbytes])
            p += 1 + rowbytes
            for i in range(rowbytes):
                a = line[i - bpp] if i >= bpp else 0
                b = prev[i]
                c = prev[i - bpp] if i >= bpp else 0
                if f == 1:
                    line[i] = (line[i] + a) & 255
                elif f == 2:
                    line[i] = (line[i] + b) & 255
                elif f == 3:
                    line[i] = (line[i] + (a + b) // 2) & 255
                elif f == 4:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    pr = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                    line[i] = (line[i] + pr) & 255
            prev = line
            if bd == 16:
                s = np.frombuffer(bytes(line), ">u2").astype(np.uint32)
            elif bd == 8:
                s = np.frombuffer(bytes(line), np.uint8).astype(np.uint32)
            else:
                bits = np.unpackbits(np.frombuffer(bytes(line), np.uint8))
                s = bits.reshape(-1, bd).dot(1 << np.arange(bd - 1, -1, -1)).astype(np.uint32)
            s = s[:pw * nch].reshape(pw, nch)
            out[y0 + y * dy, x0::dx][:pw] = s
    info["samples"] = out
    return info


def present_like_png_reader(info):
    """(imgtype name, maxval, HxWxC samples) as the reference's PNG reader
    hands the image to the library (src/imagew-png.c:312-343)."""
    s, ct, bd, trns = info["samples"], info["color_type"], info["bit_depth"], info["trns"]
    if ct == 3:
        pal = info["plte"][s[..., 0]]
        if trns is not None:
            alpha = np.full(len(info["plte"]), 255, np.uint32)
            t = np.frombuffer(trns, np.uint8)
            alpha[:len(t)] = t
            return "RGB_ALPHA", 255, np.concatenate([pal, alpha[s[..., 0]][..., None]], axis=2)
        return "RGB", 255, pal.astype(np.uint32)
    maxv = (1 << bd) - 1
    if trns is not None and ct in (0, 2):
        # png_set_tRNS_to_alpha: gray below 8 bits is expanded to 8 bits.
        if ct == 0:
            key = struct.unpack(">H", trns[:2])[0]
            a = np.where(s[..., 0] == key, 0, max(maxv, 255))
            g = s[..., 0]
            if bd < 8:
                g = g * (255 // maxv)
                maxv = 255
            return "GRAYSCALE_ALPHA", maxv, np.stack([g, a], axis=2)
        key = struct.unpack(">HHH", trns[:6])
        m = (s[..., 0] == key[0]) & (s[..., 1] == key[1]) & (s[..., 2] == key[2])
        a = np.where(m, 0, maxv)
        return "RGB_ALPHA", maxv, np.concatenate([s, a[..., None]], axis=2)
    name = {0: "GRAYSCALE", 2: "RGB", 4: "GRAYSCALE_ALPHA", 6: "RGB_ALPHA"}[ct]
    return name, maxv, s


def write_pam(path, tupl, maxval, s):
    h, w, c = s.shape
    hdr = "P7\nWIDTH %d\nHEIGHT %d\nDEPTH %d\nMAXVAL %d\nTUPLTYPE %s\nENDHDR\n" % (w, h, c, maxval, tupl)
    body = s.astype(">u2" if maxval > 255 else np.uint8).tobytes()
    with open(path, "wb") as f:
        f.write(hdr.encode() + body)


def canonical_from_png(info):
    """Expected PNG -> (H,W,4) uint32 at its sample depth; alpha-0 colour = 0."""
    s, ct, bd, trns = info["samples"], info["color_type"], info["bit_depth"], info["trns"]
    if ct == 3:
        rgb = info["plte"][s[..., 0]].astype(np.uint32)
        alpha = np.full(len(info["plte"]), 255, np.uint32)
        if trns is not None:
            t = np.frombuffer(trns, np.uint8)
            alpha[:len(t)] = t
        a = alpha[s[..., 0]]
        depth = 8
    else:
        depth = 16 if bd == 16 else 8
        full = (1 << depth) - 1
        scale = full // ((1 << bd) - 1)
        if ct in (0, 4):
            g = s[..., 0]
            rgb = np.stack([g, g, g], axis=2)
        else:
            rgb = s[..., :3]
        if ct in (4, 6):
            a = s[..., -1]
        else:
            a = np.full(s.shape[:2], (1 << bd) - 1, np.uint32)
            if trns is not None:
                if ct == 0:
                    key = struct.unpack(">H", trns[:2])[0]
                    a = np.where(s[..., 0] == key, 0, a)
                else:
                    key = struct.unpack(">HHH", trns[:6])
                    m = (s[..., 0] == key[0]) & (s[..., 1] == key[1]) & (s[..., 2] == key[2])
                    a = np.where(m, 0, a)
        rgb = rgb * scale
        a = a * scale
    out = np.concatenate([rgb, a[..., None]], axis=2).astype(np.uint32)
    out[out[..., 3] == 0, :3] = 0
    return depth, out


# (case name, source image, arguments) from the reference's tests/runtest.
SCALE = ["-width", "35", "-height", "35"]
SCALE2 = ["-width", "24", "-height", "24"]
SMALL = ["-width", "15", "-height", "15"]


def cases():
    c = []
    for f in ["g1", "g2", "g4", "g8", "p1", "p1t", "p2", "p2t", "p4", "p4t", "p8", "p8t", "rgb8", "g8a", "rgb8a"]:
        c.append(("png-" + f, f, SCALE + ["-filter", "catrom"]))
    for f in ["g16", "rgb16", "g16a", "rgb16a"]:
        c.append(("png-" + f, f, SCALE + ["-filter", "catrom", "-depth", "16"]))
    for f in ["g1t", "g2t", "g4t", "g8t", "rgb8t"]:
        c.append(("png-%sns" % f, f, []))
    for f in ["g16t", "rgb16t"]:
        c.append(("png-%sns" % f, f, ["-depth", "16"]))
    c += [
        ("ccgray-4", "g8", SMALL + ["-filter", "bspline", "-cc", "4", "-dither", "o"]),
        ("ccgray-16", "g8", SMALL + ["-filter", "bspline", "-cc", "16", "-dither", "o"]),
        ("ccgray-17", "g8", SMALL + ["-filter", "bspline", "-cc", "17", "-dither", "o"]),
        ("ccgraya-4", "g8a", SMALL + ["-filter", "bspline", "-cc", "4", "-dither", "o"]),
        ("ccgray-10", "rgb8", SMALL + ["-filter", "gaussian", "-cc", "10", "-dither", "o", "-grayscale"]),
        ("condgray", "g4", SMALL + ["-filter", "mitchell", "-cc", "7", "-ccgray", "6", "-ccred", "5", "-condgrayscale"]),
        ("bkgd", "g2t", SCALE + ["-filter", "catrom", "-bkgd", "e42"]),
        ("bkgd2", "rgb8a", SCALE + ["-filter", "catrom", "-bkgd", "e42,0f5", "-checkersize", "6", "-checkerorigin", "1,3"]),
        ("bkgd3", "rgb8a", SCALE + ["-filter", "catrom", "-bkgd", "e42d,00ff5550", "-checkersize", "5",
                                    "-checkerorigin", "2,5", "-edge", "t", "-translate", "4,3"]),
        ("bt-g1", "bt-gray", SCALE + ["-filter", "lanczos"]),
        ("bt-w1", "bt-white", SCALE + ["-filter", "lanczos"]),
        ("bt-b1", "g1t", SCALE + ["-filter", "lanczos"]),
        ("bt-g2", "bt-gray", SCALE + ["-filter", "lanczos", "-bkgd", "111"]),
    ]
    for f in ["auto", "mix", "triangle", "gaussian", "lanczos"]:
        c.append(("ds-" + f, "rings1", SCALE + ["-filter", f]))
    c += [
        ("ds-linearinterp", "rings1", ["-w", "35", "-h", "35", "-filter", "triangle", "-blur", "x1"]),
        ("ds-box", "rgb8", ["-w", "20", "-h", "20", "-filter", "box", "-translate", "0.0000001,0", "-quiet"]),
        ("ds-boxavg", "rgb8", ["-w", "20", "-h", "20", "-filter", "boxavg"]),
        ("ds-nearest", "4x4", ["-w", "2", "-h", "2", "-filter", "nearest", "-translate", "0.0000001,0", "-quiet"]),
        ("cc-2", "4x4", SCALE + ["-filter", "catrom", "-cc", "2"]),
        ("cc-6", "4x4", SCALE + ["-filter", "catrom", "-cc", "6"]),
        ("cc-mixed", "4x4", SCALE + ["-filter", "catrom", "-cc", "3,10,5"]),
        ("depth-16", "4x4", SCALE + ["-filter", "catrom", "-depth", "16"]),
    ]
    for f in ["auto", "nearest", "mix", "box", "triangle", "quadratic", "gaussian", "hermite",
              "catrom", "mitchell", "bspline", "lanczos2", "lanczos", "lanczos8", "blackman", "hanning"]:
        c.append(("us-" + f, "4x4", SCALE + ["-filter", f]))
    c += [
        ("us-cubic01", "4x4", SCALE + ["-filter", "cubic0,1", "-interlace"]),
        ("us-mixed", "4x4", SCALE + ["-filterx", "catrom", "-filtery", "nearest"]),
        ("blur", "4x4", SCALE + ["-filter", "catrom", "-blur", "1.5"]),
        ("edge-s", "4x4", SCALE + ["-filter", "lanczos", "-edge", "s"]),
        ("edge-r", "4x4", SCALE + ["-filter", "lanczos", "-edge", "r"]),
        ("edge-t", "4x4", SCALE + ["-filter", "mitchell", "-edgex", "t", "-edgey", "r"]),
        ("edge-t2", "p8t", SMALL + ["-filter", "hanning", "-edge", "t", "-translate", "s3,3"]),
        ("edge-t3", "rgb8", SMALL + ["-filter", "lanczos5", "-edge", "t", "-translate", "-4,-3", "-grayscale",
                                    "-bkgd", "876,554433", "-checkersize", "3"]),
        ("edge-t4", "rgb8", SMALL + ["-filter", "cubic0,2.2", "-edge", "t", "-translate", "3,4", "-offsetrb", ".3",
                                    "-bkgd", "965"]),
        ("cs-linear", "4x4", SCALE + ["-filter", "catrom", "-cs", "linear"]),
        ("cs-gamma15", "4x4", ["-height", "x8.75", "-filter", "catrom", "-cs", "gamma1.5"]),
        ("cs-srgbs", "4x4", ["-width", "x8.75", "-filter", "catrom", "-cs", "srgbs"]),
        ("nocslabel", "4x4", SCALE + ["-filter", "catrom", "-nocslabel"]),
        ("inputcs", "4x4", SCALE + ["-filter", "catrom", "-inputcs", "linear", "-cs", "srgb"]),
        ("nogamma", "4x4", SCALE + ["-filter", "catrom", "-nogamma"]),
        ("intclamp", "4x4", SCALE + ["-filter", "lanczos", "-intclamp"]),
        ("grayscale", "4x4", SCALE + ["-filter", "catrom", "-grayscale"]),
        ("grayscale-c", "4x4", SCALE + ["-filter", "catrom", "-nogamma", "-grayscaleformula", "c"]),
        ("grayscale-w", "4x4", SCALE + ["-filter", "quadratic", "-gsf", "w3,7,1"]),
        ("grayscale-v", "4x4", SCALE + ["-blur", "1", "-gsf", "v.4,.1,.4", "-grayscale", "-nowarn"]),
        ("offset", "4x4", SCALE + ["-filter", "mix", "-offsetred", ".333", "-offsetgreen", "-0.2",
                                   "-offsetblue", "-1.5", "-edge", "r"]),
        ("offsetv", "rgb8a", SCALE + ["-filter", "mix", "-offsetvred", ".333", "-offsetvgreen", "-0.2",
                                     "-offsetvblue", "-1.5", "-edge", "r", "-nowarn"]),
        ("offsetrb", "g2", SCALE + ["-filter", "catrom", "-offsetrb", ".333", "-offsetvrb", "-0.6", "-edge", "r"]),
    ]
    for d in ["f", "o", "halftone", "sierra", "sierra2", "sierralite", "jjn", "burkes", "atkinson", "r", "r2"]:
        c.append(("dither-" + d, "4x4", SCALE + ["-filter", "catrom", "-cc", "3", "-dither", d]))
    c += [
        ("dither-gray", "4x4", SCALE + ["-filter", "catrom", "-cc", "2", "-grayscale", "-dither", "f"]),
        ("imgsize1", "p8t", ["-S", "35,35", "-translate", "5.2,5.9", "-imagesize", "24,25.5", "-edge", "t",
                             "-bkgd", "987,654", "-filter", "lanczos4"]),
        ("dens-2", "rgb8x2", SCALE + ["-filter", "mix"]),
        ("dens-1", "rgb8x1", SCALE + ["-filter", "mix", "-cs", "rec709"]),
        ("dens-ix", "25x20", ["-w", "34"]),
        ("dens-iy", "25x20", ["-h", "27"]),
        ("dens-ixy", "25x20", ["-w", "34", "-h", "27"]),
        ("dens-imgsize", "25x20", ["-S", "34,34", "-translate", "1,2.5", "-imagesize", "31,25", "-edge", "t",
                                   "-bkgd", "808b"]),
        ("orient1", "25x20", ["-reorient", "transverse"]),
        ("opt-02col", "rgb8a", ["-width", "2", "-height", "1", "-filter", "mix"]),
        ("opt-04col", "rgb8a", ["-width", "2", "-height", "2", "-filter", "mix"]),
        ("opt-12col", "rgb8", ["-width", "4", "-height", "3", "-filter", "mix"]),
        ("opt-20col", "rgb8a", ["-width", "5", "-height", "4", "-filter", "mix"]),
        ("crop-1", "p8t", ["-width", "20", "-crop", "3,12,18,9"]),
        ("noopt-g", "g8a", ["-width", "15", "-noopt", "grayscale", "-filter", "mix"]),
        ("noopt-p", "rgb8", ["-width", "15", "-noopt", "palette", "-filter", "mix"]),
        ("noopt-sa", "p4t", ["-crop", "0,0,6,6", "-width", "15", "-noopt", "stripalpha"]),
        ("noopt-bt", "g8a", ["-ccalpha", "2", "-dither", "f", "-width", "15", "-noopt", "binarytrns", "-filter", "mix"]),
        ("noopt-r8", "g8", ["-depth", "16", "-crop", "17,18,-1", "-noopt", "reduceto8"]),
        ("bkgdlabel1", "rgb8", ["-width", "15", "-noopt", "palette", "-bkgdlabel", "2380ff"]),
        ("bkgdlabel5", "rgb8", ["-depth", "16", "-bkgdlabel", "30398707ddd5"]),
        ("bkgdlabel6", "rgb8", ["-depth", "16", "-bkgdlabel", "30308787dddd"]),
        ("bkgdlabel7", "rgb8", ["-width", "15", "-bkgdlabel", "3e4", "-grayscale"]),
        ("bkgdlabel8", "g4t", ["-bkgdlabel", "444"]),
        ("bkgdlabel9", "g4t", ["-bkgdlabel", "454545"]),
        ("neg1", "rgb8a", SMALL + ["-negate"]),
    ]
    return c


def main():
    os.makedirs(os.path.join(HERE, "srcimg"), exist_ok=True)
    all_cases = cases()
    used = sorted({src for _, src, _ in all_cases})
    for src in used:
        info = read_png(os.path.join(REF, "srcimg", src + ".png"))
        tupl, maxval, s = present_like_png_reader(info)
        write_pam(os.path.join(HERE, "srcimg", src + ".pam"), tupl, maxval, s)
    exp = {}
    depth = {}
    for name, _, _ in all_cases:
        info = read_png(os.path.join(REF, "expected", name + ".png"))
        d, arr = canonical_from_png(info)
        exp[name] = arr.astype(np.uint16)
        depth[name] = d
    np.savez_compressed(os.path.join(HERE, "expected.npz"), **exp)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump([{"name": n, "src": s, "args": a, "depth": depth[n]} for n, s, a in all_cases], f, indent=0)
    print("wrote %d cases, %d inputs" % (len(all_cases), len(used)))


if __name__ == "__main__":
    sys.exit(main())
