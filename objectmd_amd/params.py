"""Readers of Object-MD parameter files (tables/sw.silicon,
tables/tersoff.silicon): "key value..." lines, '#' comments
(src/omd/param.cpp:106-139)."""


def read_params(path):
    out = {}
    with open(path) as fl:
        for line in fl:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            out[t[0]] = t[1:]
    return out


def sw_params(path, z1=14.0, z2=14.0):
    """The 19 values of omdg_set_sw from a tables/sw.<material> file (keys of
    sw.cpp:35-60); z1, z2 are the atomic numbers of the ZBL core."""
    p = read_params(path)
    g = lambda k, i=0: float(p[k][i])
    usezbl = p.get("zbl", ["no"])[0] in ("yes", "true", "1")
    return [g("sw.epsilon"), g("sw.A"), g("sw.B"), g("sw.p"), g("sw.q"), g("sw.sigma"),
            g("sw.a"), g("sw.lambda"), g("sw.gamma"), g("sw.invers_cos0"),
            1.0 if usezbl else 0.0, g("spline.range", 0), g("spline.range", 1),
            g("spline.coeff", 0), g("spline.coeff", 1), g("spline.coeff", 2),
            g("spline.coeff", 3), z1, z2]


def tersoff_params(path, R=None, D=None):
    """The 13 values of omdg_set_tersoff from a tables/tersoff.<material>
    file (keys of tersoff.h:146-180); R and D may be overridden."""
    p = read_params(path)
    g = lambda k: float(p[k][0])
    return [g("tersoff.A"), g("tersoff.B"), g("tersoff.lambda_1"), g("tersoff.lambda_2"),
            g("tersoff.lambda_3"), g("tersoff.alpha"), g("tersoff.beta"), g("tersoff.n"),
            g("tersoff.c"), g("tersoff.d"), g("tersoff.h"),
            g("tersoff.R") if R is None else R, g("tersoff.D") if D is None else D]
