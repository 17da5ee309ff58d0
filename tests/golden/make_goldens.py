"""Writes reference_goldens.json: golden vectors transcribed from the
reference's own tests (Stencils.jl v0.3.6 test/array.jl, test/stencils.jl).

Each case is data only: the input, the stencil / boundary / padding, the
function and the expected output, with the file:line it comes from.  Julia
arrays are written column-major as nested lists indexed [i][j] = A[i+1, j+1].
`tol` is null for exact (==) comparisons and the absolute tolerance the
reference's `≈` implies otherwise (its goldens are printed to ~8-11 digits).

    python tests/golden/make_goldens.py
"""
import json
from pathlib import Path

r = [[float(i * j) for j in range(100, 106)] for i in range(1, 6)]  # (1.0:5.0) * (100.0:105.0)'

init = [[0, 0, 0, 1, 1, 1],
        [1, 0, 1, 1, 0, 1],
        [0, 1, 1, 1, 1, 1],
        [0, 1, 0, 0, 1, 0],
        [0, 0, 0, 0, 1, 1],
        [0, 1, 0, 1, 1, 0]]
win = [[0, 1, 0, 0, 1],
       [0, 0, 1, 0, 0],
       [0, 0, 0, 1, 1],
       [0, 0, 1, 0, 1],
       [1, 0, 1, 0, 1]]


def madd(a, b):
    return [[x + y for x, y in zip(ra, rb)] for ra, rb in zip(a, b)]


CASES = [
    # ---------------------------------------------------------- test/array.jl
    dict(name="remove_mean_2d", src="test/array.jl:123-145", input=r, dtype="f64",
         stencil=["Window", 1, 2], boundary=["Remove", 0.0], paddings=["Conditional", "Halo_out"],
         f="mean", tol=1e-8,
         expected=[[67.0, 101.0, 102.0, 103.0, 104.0, 69.66666666667],
                   [134.0, 202.0, 204.0, 206.0, 208.0, 139.33333333333],
                   [201.0, 303.0, 306.0, 309.0, 312.0, 209.0],
                   [268.0, 404.0, 408.0, 412.0, 416.0, 278.66666666667],
                   [201.0, 303.0, 306.0, 309.0, 312.0, 209.0]]),
    dict(name="use_halo_in_mean_2d", src="test/array.jl:127,135,147", input=r, dtype="f64",
         stencil=["Window", 1, 2], boundary=["Use"], paddings=["Halo_in"], f="mean", tol=None,
         expected=[row[1:-1] for row in r[1:-1]]),
    dict(name="remove_sum_2d_interior", src="test/array.jl:150-161", input=r, dtype="f64",
         stencil=["Window", 1, 2], boundary=["Remove", 0.0], paddings=["Conditional", "Halo_out"],
         f="sum", tol=None, interior=1,
         expected=[[1818.0, 1836.0, 1854.0, 1872.0],
                   [2727.0, 2754.0, 2781.0, 2808.0],
                   [3636.0, 3672.0, 3708.0, 3744.0]]),
    dict(name="use_halo_in_sum_2d", src="test/array.jl:152,157-161", input=r, dtype="f64",
         stencil=["Window", 1, 2], boundary=["Use"], paddings=["Halo_in"], f="sum", tol=None,
         expected=[[1818.0, 1836.0, 1854.0, 1872.0],
                   [2727.0, 2754.0, 2781.0, 2808.0],
                   [3636.0, 3672.0, 3708.0, 3744.0]]),
    dict(name="wrap_mean_1d", src="test/array.jl:165-181", input=[1.0, 2.0, 3.0, 4.0, 5.0],
         dtype="f64", stencil=["Window", 1, 1], boundary=["Wrap"],
         paddings=["Conditional", "Halo_out"], f="mean", tol=None,
         expected=[2.6666666666666665, 2.0, 3.0, 4.0, 3.3333333333333335]),
    dict(name="wrap_mean_1d_halo_in", src="test/array.jl:170,177,182", input=[1.0, 2.0, 3.0, 4.0, 5.0],
         dtype="f64", stencil=["Window", 1, 1], boundary=["Wrap"], paddings=["Halo_in"],
         f="mean", tol=None, expected=[3.0, 3.0, 3.0]),
    dict(name="wrap_mean_2d", src="test/array.jl:184-206", input=r, dtype="f64",
         stencil=["Window", 1, 2], boundary=["Wrap"], paddings=["Conditional", "Halo_out"],
         f="mean", tol=1e-7,
         expected=[[272.0, 269.33333333, 272.0, 274.66666666, 277.33333333, 274.66666666],
                   [204.0, 202.0, 204.0, 206.0, 208.0, 206.0],
                   [306.0, 303.0, 306.0, 309.0, 312.0, 309.0],
                   [408.0, 404.0, 408.0, 412.0, 416.0, 412.0],
                   [340.0, 336.66666666, 340.0, 343.33333333, 346.66666666, 343.33333333]]),
    dict(name="reflect_mean_1d", src="test/array.jl:212-228", input=[1.0, 2.0, 3.0, 4.0, 5.0],
         dtype="f64", stencil=["Window", 1, 1], boundary=["Reflect"],
         paddings=["Conditional", "Halo_out"], f="mean", tol=1e-7,
         expected=[1.66666666, 2.0, 3.0, 4.0, 4.33333333]),
    dict(name="reflect_mean_1d_halo_in", src="test/array.jl:217,224,229",
         input=[1.0, 2.0, 3.0, 4.0, 5.0], dtype="f64", stencil=["Window", 1, 1],
         boundary=["Reflect"], paddings=["Halo_in"], f="mean", tol=1e-7,
         expected=[2.66666666, 3.0, 3.33333333]),
    dict(name="reflect_mean_2d", src="test/array.jl:231-253", input=r, dtype="f64",
         stencil=["Window", 1, 2], boundary=["Reflect"], paddings=["Conditional", "Halo_out"],
         f="mean", tol=1e-7,
         expected=[[167.77777777, 168.33333333, 170.0, 171.66666666, 173.33333333, 173.888888888],
                   [201.33333333, 202.0, 204.0, 206.0, 208.0, 208.666666666],
                   [302.0, 303.0, 306.0, 309.0, 312.0, 313.0],
                   [402.66666666, 404.0, 408.0, 412.0, 416.0, 417.333333333],
                   [436.22222222, 437.66666666, 442.0, 446.33333333, 450.66666666, 452.111111111]]),
    # --------------------------------------------------------- test/stencils.jl
    # mapstencil(s -> s.n + s.w + center(s), StencilArray(win, h1)) == M + win:
    # the sum over the named offsets n=(-1,0), w=(1,0) plus the centre.
    dict(name="named_n_plus_w_plus_center", src="test/stencils.jl:194,205-213", input=win,
         dtype="i64", stencil=["Positional", [[-1, 0], [1, 0], [0, 0]]], boundary=["Remove", 0],
         paddings=["Conditional"], f="sum", tol=None,
         expected=madd([[0, 0, 1, 0, 0],
                        [0, 1, 0, 1, 2],
                        [0, 0, 2, 0, 1],
                        [1, 0, 1, 1, 2],
                        [0, 0, 1, 0, 1]], win)),
    # NamedStencil(Cardinal(1)) names (E, S, N, W) = ((0,-1), (-1,0), (1,0), (0,1)); s.W + s.S
    dict(name="cardinal_w_plus_s", src="test/stencils.jl:219-228", input=win, dtype="i64",
         stencil=["Positional", [[0, 1], [-1, 0]]], boundary=["Remove", 0],
         paddings=["Conditional"], f="sum", tol=None,
         expected=[[1, 0, 0, 1, 0],
                   [0, 2, 0, 0, 1],
                   [0, 0, 2, 1, 0],
                   [0, 1, 0, 2, 1],
                   [0, 1, 1, 1, 1]]),
    # NamedStencil(Ordinal(1)) names (SE, NE, SW, NW) = ((-1,-1), (1,-1), (-1,1), (1,1)); s.NE + s.NW
    dict(name="ordinal_ne_plus_nw", src="test/stencils.jl:230-239", input=win, dtype="i64",
         stencil=["Positional", [[1, -1], [1, 1]]], boundary=["Remove", 0],
         paddings=["Conditional"], f="sum", tol=None,
         expected=[[0, 1, 0, 1, 0],
                   [0, 0, 1, 1, 1],
                   [0, 1, 0, 2, 0],
                   [0, 2, 0, 2, 0],
                   [0, 0, 0, 0, 0]]),
]

# Single-cell known answers: sum(neighbours) of a stencil at one index.
POINT_CASES = [
    dict(name="vonneumann_init_2_2", src="test/stencils.jl:59-72", input=init, stencil=["VonNeumann", 1, 2],
         index=[1, 1], neighbors=[1, 0, 1, 1], sum=3),
    dict(name="annulus_init_2_2", src="test/stencils.jl:79-93", input=init, stencil=["Annulus", 1, 0, 2],
         index=[1, 1], neighbors=[0, 1, 0, 0, 1, 0, 1, 1], sum=4),
    dict(name="ordinal_init_2_2", src="test/stencils.jl:99-112", input=init, stencil=["Ordinal", 1, 2],
         index=[1, 1], neighbors=[0, 0, 0, 1], sum=1),
    dict(name="cardinal_init_2_2", src="test/stencils.jl:117-130", input=init, stencil=["Cardinal", 1, 2],
         index=[1, 1], neighbors=[1, 0, 1, 1], sum=3),
    dict(name="positional_h1_win_3_3", src="test/stencils.jl:141-148", input=win,
         stencil=["Positional", [[-1, -1], [2, -2], [2, 2], [-1, 2], [0, 0]]],
         index=[2, 2], neighbors=[0, 1, 1, 0, 0], sum=2),
    dict(name="positional_h2_win_3_3", src="test/stencils.jl:150-156", input=win,
         stencil=["Positional", [[-1, -1], [0, -1], [1, -1], [0, 0]]],
         index=[2, 2], neighbors=[0, 0, 0, 0], sum=0),
    dict(name="rectangle_h1_win_3_3", src="test/stencils.jl:165-174", input=win,
         stencil=["Rectangle", [[-1, 0], [-2, 1]]], index=[2, 2],
         neighbors=[0, 0, 0, 0, 1, 0, 0, 1], sum=2),
    dict(name="rectangle_h3_3d_2_2_1", src="test/stencils.jl:176-184", input3=[win, win, win],
         stencil=["Rectangle", [[-1, 0], [0, 1], [1, 1]]], index=[1, 1, 0],
         neighbors=[1, 0, 0, 1], sum=2),
]

# Offset tables (test/stencils.jl) and boundary index maps (test/array.jl:5-14).
OFFSETS = [
    dict(src="test/stencils.jl:28", stencil=["Moore", 1, 2],
         offsets=[[-1, -1], [0, -1], [1, -1], [-1, 0], [1, 0], [-1, 1], [0, 1], [1, 1]]),
    dict(src="test/stencils.jl:48-49", stencil=["Window", 1, 2],
         offsets=[[-1, -1], [0, -1], [1, -1], [-1, 0], [0, 0], [1, 0], [-1, 1], [0, 1], [1, 1]]),
    dict(src="test/stencils.jl:63", stencil=["VonNeumann", 1, 2],
         offsets=[[0, -1], [-1, 0], [1, 0], [0, 1]]),
    dict(src="test/stencils.jl:74-76", stencil=["VonNeumann", 2, 2],
         offsets=[[0, -2], [-1, -1], [0, -1], [1, -1], [-2, 0], [-1, 0], [1, 0], [2, 0],
                  [-1, 1], [0, 1], [1, 1], [0, 2]]),
    dict(src="test/stencils.jl:83-84", stencil=["Annulus", 1, 0, 2],
         offsets=[[-1, -1], [0, -1], [1, -1], [-1, 0], [1, 0], [-1, 1], [0, 1], [1, 1]]),
    dict(src="test/stencils.jl:95-96", stencil=["Annulus", 2, 1, 2],
         offsets=[[-1, -2], [0, -2], [1, -2], [-2, -1], [2, -1], [-2, 0], [2, 0], [-2, 1],
                  [2, 1], [-1, 2], [0, 2], [1, 2]]),
    dict(src="test/stencils.jl:103", stencil=["Ordinal", 1, 2],
         offsets=[[-1, -1], [1, -1], [-1, 1], [1, 1]]),
    dict(src="test/stencils.jl:114", stencil=["Ordinal", 2, 2],
         offsets=[[-2, -2], [2, -2], [-2, 2], [2, 2]]),
    dict(src="test/stencils.jl:121", stencil=["Cardinal", 1, 2],
         offsets=[[0, -1], [-1, 0], [1, 0], [0, 1]]),
    dict(src="test/stencils.jl:132", stencil=["Cardinal", 2, 2],
         offsets=[[0, -2], [-2, 0], [2, 0], [0, 2]]),
]
INDICES = [
    # indices(A, (1, 1)) on a 4x4 Moore(1) array, 1-based as in the test
    dict(src="test/array.jl:6-7", boundary="Remove",
         expected=[[0, 0], [1, 0], [2, 0], [0, 1], [2, 1], [0, 2], [1, 2], [2, 2]]),
    dict(src="test/array.jl:8-9", boundary="Wrap",
         expected=[[4, 4], [1, 4], [2, 4], [4, 1], [2, 1], [4, 2], [1, 2], [2, 2]]),
    dict(src="test/array.jl:10-11", boundary="Reflect",
         expected=[[2, 2], [1, 2], [2, 2], [2, 1], [2, 1], [2, 2], [1, 2], [2, 2]]),
]
KERNELPRODUCT = [
    # Kernel(Window{1,2}(1:9, 5), reshape(1:9, 3, 3)): sum((1:9).^2)
    dict(src="test/stencils.jl:276-277", hood=list(range(1, 10)), weights=list(range(1, 10)),
         stencil=["Window", 1, 2], expected=285),
    # Positional((0,-1),(-1,0),(1,0),(0,1)) on reshape(1:9,3,3) at (2,2), kernel 1:4
    dict(src="test/stencils.jl:296-303", window=[[1, 4, 7], [2, 5, 8], [3, 6, 9]],
         stencil=["Positional", [[0, -1], [-1, 0], [1, 0], [0, 1]]], weights=[1, 2, 3, 4],
         index=[1, 1], expected=60),
]

if __name__ == "__main__":
    out = dict(reference="Stencils.jl v0.3.6", cases=CASES, points=POINT_CASES,
               offsets=OFFSETS, indices=INDICES, kernelproduct=KERNELPRODUCT)
    p = Path(__file__).with_name("reference_goldens.json")
    p.write_text(json.dumps(out, indent=1))
    print(f"wrote {p}")
