"""Host model of the single-step row tile's LDS plane (cfdarcov2_amd/csrc/row.hpp,
row_tile): every wave role and lane of a workgroup, for tiles of 4, 8 and 14
output waves and z extents with one or several rows per wave.

Per plane the tile writes its p row slots (phase 1), reads the y-neighbours
and the halo rows' inner neighbours (phase 2), writes its v' slots and reads
them again (phase 3), with a barrier between phases.  The model checks, for
every lane whose result is stored, that each slot it reads lies inside the
(TY + 2) x 64 plane and was written in the preceding phase by exactly the
lane that holds that row and column, and that no two lanes write one slot
in a phase.  This is the index arithmetic of the 14-row FP64 tile too (the
element type only scales the byte offsets)."""
import pytest


def rows_per_wave(Z):
    return 1 if Z >= 256 else 256 // Z


def lanes(TY, Z, Y, y0):
    RW = rows_per_wave(Z)
    zq = Z // 4
    out = []
    for wv in range(TY + 2):
        is_out = wv < TY
        lower = wv == TY
        for lane in range(64):
            if is_out:
                c = 4 * lane
                r, zc, span = c // Z, c % Z, c < RW * Z
            else:
                r, zc, span = 0, 4 * lane, lane < zq
            y = (y0 + wv * RW + r) if is_out else (y0 - 1 if lower else y0 + TY * RW)
            real = span and 0 <= y < Y
            own = (1 + wv * RW) * zq + lane if is_out else (lane if lower else (1 + TY * RW) * zq + lane)
            ilo = own if y == 0 else own - zq
            ihi = own if y == Y - 1 else own + zq
            iin = own + zq if lower else own - zq
            out.append(dict(wv=wv, lane=lane, is_out=is_out, lower=lower, span=span, y=y, zc=zc, real=real,
                            own=own, ilo=ilo, ihi=ihi, iin=iin))
    return out


@pytest.mark.parametrize("TY", [4, 8, 14])
@pytest.mark.parametrize("Z", [4, 8, 12, 16, 32, 64, 100, 128, 256])
def test_row_tile_lds_slots(TY, Z):
    RW = rows_per_wave(Z)
    rows = TY * RW
    size = (TY + 2) * 64
    for Y in (1, 2, rows - 1, rows, rows + 3, 3 * rows + 1):
        if Y < 1:
            continue
        for y0 in range(0, Y, rows):
            L = lanes(TY, Z, Y, y0)
            waves_live = {w for w in range(TY) if any(l["real"] for l in L if l["wv"] == w)}
            # phase 1: p slots; output waves of live waves write where span, halo waves where real
            wr1 = {}
            for l in L:
                w = (l["is_out"] and l["wv"] in waves_live and l["span"]) or (not l["is_out"] and l["real"])
                if w:
                    assert 0 <= l["own"] < size
                    assert l["own"] not in wr1, (TY, Z, Y, y0, l)
                    wr1[l["own"]] = (l["y"], l["zc"])
            # phase 2 reads: stored lanes read their y-neighbours' p; halo lanes read the inner row
            for l in L:
                if l["is_out"] and l["real"]:
                    for key, dy in (("ilo", -1), ("ihi", 1)):
                        s = l[key]
                        want_y = l["y"] if (l["y"] + dy < 0 or l["y"] + dy >= Y) else l["y"] + dy
                        assert 0 <= s < size and s in wr1, (TY, Z, Y, y0, key, l)
                        assert wr1[s] == (want_y, l["zc"]), (TY, Z, Y, y0, key, l, wr1[s])
                if not l["is_out"] and l["real"]:
                    s = l["iin"]
                    inner = l["y"] + 1 if l["lower"] else l["y"] - 1
                    if 0 <= inner < Y:
                        assert s in wr1 and wr1[s] == (inner, l["zc"]), (TY, Z, Y, y0, l)
            # phase 2 writes v' slots (same places as phase 1's); phase 3 reads them like phase 2's p reads
            wr2 = {}
            for l in L:
                w = (l["is_out"] and l["wv"] in waves_live and l["span"]) or (not l["is_out"] and l["real"])
                if w:
                    assert l["own"] not in wr2
                    wr2[l["own"]] = (l["y"], l["zc"])
            assert wr2 == wr1
