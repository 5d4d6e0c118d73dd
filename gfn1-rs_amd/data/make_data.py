#!/usr/bin/env python3
"""Regenerate the model data files of this directory from their upstream sources.

Run once, by hand, where the upstream sources are available; the outputs are
committed and nothing reads the sources at run time.

Outputs
-------
param_gfn1-xtb.txt
    The official GFN1-xTB parameter file (grimme-lab/xtb, LGPL-3.0-or-later),
    copied byte for byte.  See NOTICE.
sto_ng.txt
    Stewart's least-squares STO-nG expansions (R. F. Stewart, J. Chem. Phys.
    52, 431 (1970)): for each contraction length ng = 1..6 and each Slater
    type (1s 2s 3s 4s 5s 2p 3p 4p 5p 3d 4d 4f 5f 5g ...), the ng exponents
    (for zeta = 1) and contraction coefficients, plus the separate 6s/6p
    STO-6G expansions.  Format: one record per line,
    ``<tag> <ng> <alpha_1..alpha_ng> <coeff_1..coeff_ng>`` where ``tag`` is
    the 1-based Slater-type index 1..15 or ``6s``/``6p``.
d3_reference.bin
    DFT-D3 reference data (s-dftd3, LGPL-3.0-or-later) for Z = 1..86 in a
    little-endian binary layout of our own:
      int32 magic 0x44334331 ('D3C1'), int32 max_elem (86), int32 max_ref (7)
      int32 number_of_references[max_elem]
      f64   reference_cn[max_elem][max_ref]
      f64   r4_over_r2_raw[max_elem]          (raw <r4>/<r2> expectation values)
      f64   c6[npair][max_ref][max_ref]       npair = max_elem(max_elem+1)/2
    c6 pair index for Za >= Zb (1-based Z): pair = Zb + Za(Za-1)/2 - 1 and
    c6[pair][ref_b][ref_a] is C6 between reference ref_a of Za and ref_b of Zb.
"""
import os
import re
import shutil
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
UPSTREAM = os.environ.get("GFN1_UPSTREAM", "/root/reference")
MAX_ELEM_OUT = 86


def strip_bang_comments(text):
    return "\n".join(line.split("!")[0] for line in text.splitlines())


def numbers(text):
    out = []
    for tok in re.findall(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][-+]?\d+)?", text):
        out.append(float(tok.replace("d", "e").replace("D", "e")))
    return out


def bracket_body(text, open_idx):
    depth = 0
    for i in range(open_idx, len(text)):
        if text[i] == "[":
            depth += 1
        elif text[i] == "]":
            depth -= 1
            if depth == 0:
                return text[open_idx + 1:i], i
    raise ValueError("unbalanced bracket")


def make_d3():
    src = os.path.join(UPSTREAM, "third_party/simple-dftd3/src/dftd3")
    ref = strip_bang_comments(open(os.path.join(src, "reference.f90")).read())
    r4 = strip_bang_comments(open(os.path.join(src, "data/r4r2.f90")).read())
    max_elem = int(re.search(r"max_elem\s*=\s*(\d+)", ref).group(1))
    max_ref = int(re.search(r"max_ref\s*=\s*(\d+)", ref).group(1))

    def array_after(text, name):
        m = re.search(r"::\s*" + name + r"\b", text)
        body, _ = bracket_body(text, text.index("[", m.end()))
        # drop the Fortran kind suffix so "_wp" does not glue onto numbers
        return numbers(body.replace("_wp", " "))

    nref = [int(round(v)) for v in array_after(ref, "number_of_references")]
    refcn = array_after(ref, "reference_cn")
    # reshape([...], [max_ref, max_elem]) -- the trailing shape literal is a
    # second bracket; array_after stops at the first closing bracket.
    assert len(nref) == max_elem, len(nref)
    assert len(refcn) == max_ref * max_elem, len(refcn)

    total = max_ref * max_ref * max_elem * (max_elem + 1) // 2
    c6 = [None] * total
    for m in re.finditer(r"c6ab_view\((\d+)\s*:\s*(\d+)\)\s*=\s*\[", ref):
        lo, hi = int(m.group(1)), int(m.group(2))
        body, _ = bracket_body(ref, m.end() - 1)
        vals = numbers(body.replace("_wp", " "))
        assert len(vals) == hi - lo + 1, (lo, hi, len(vals))
        c6[lo - 1:hi] = vals
    assert all(v is not None for v in c6), "C6 table not fully covered"

    r4raw = array_after(r4, "r4_over_r2")
    assert len(r4raw) >= MAX_ELEM_OUT

    me = MAX_ELEM_OUT
    npair = me * (me + 1) // 2
    with open(os.path.join(HERE, "d3_reference.bin"), "wb") as f:
        f.write(struct.pack("<iii", 0x44334331, me, max_ref))
        f.write(struct.pack("<%di" % me, *nref[:me]))
        f.write(struct.pack("<%dd" % (me * max_ref), *refcn[:me * max_ref]))
        f.write(struct.pack("<%dd" % me, *r4raw[:me]))
        # the upstream flat layout is pair-major and already ordered the way
        # the header describes; pairs with both Z <= 86 are the leading block
        f.write(struct.pack("<%dd" % (npair * max_ref * max_ref),
                            *c6[:npair * max_ref * max_ref]))
    print("d3_reference.bin: max_elem", me, "max_ref", max_ref, "npair", npair)


def make_sto():
    text = open(os.path.join(UPSTREAM, "src/sto.rs")).read()
    out = []
    for ng in range(1, 7):
        alphas = re.search(r"const P_ALPHA%d: \[\[f64; %d\]; NF\] = (\[.*?\]);\n" % (ng, ng),
                           text, re.S).group(1)
        a = numbers(alphas)
        if ng == 1:
            c = [1.0] * 15
        else:
            coeffs = re.search(r"const P_COEFF%d: \[\[f64; %d\]; NF\] = (\[.*?\]);\n" % (ng, ng),
                               text, re.S).group(1)
            c = numbers(coeffs)
        # numbers() also catches the 'f64; N' type tokens; strip leading junk
        a = a[-15 * ng:]
        c = c[-15 * ng:]
        assert len(a) == 15 * ng and len(c) == 15 * ng, (ng, len(a), len(c))
        for t in range(15):
            row = a[t * ng:(t + 1) * ng] + c[t * ng:(t + 1) * ng]
            out.append("%d %d " % (t + 1, ng) + " ".join("%.10e" % v for v in row))
    for tag in ("6S", "6P"):
        a = numbers(re.search(r"const P_ALPHA%s: \[f64; 6\] = (\[.*?\]);" % tag, text, re.S).group(1))[-6:]
        c = numbers(re.search(r"const P_COEFF%s: \[f64; 6\] = (\[.*?\]);" % tag, text, re.S).group(1))[-6:]
        out.append("%s 6 " % tag.lower() + " ".join("%.10e" % v for v in a + c))
    with open(os.path.join(HERE, "sto_ng.txt"), "w") as f:
        f.write("# Stewart STO-nG expansions; see make_data.py for the format.\n")
        f.write("\n".join(out) + "\n")
    print("sto_ng.txt:", len(out), "records")


def copy_params():
    shutil.copyfile(os.path.join(UPSTREAM, "third_party/xtb/param_gfn1-xtb.txt"),
                    os.path.join(HERE, "param_gfn1-xtb.txt"))


if __name__ == "__main__":
    if not os.path.isdir(UPSTREAM):
        sys.exit("upstream sources not found at %s" % UPSTREAM)
    copy_params()
    make_sto()
    make_d3()
