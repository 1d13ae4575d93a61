"""Input.txt writer for the fdtd_run tests: one header token, then the 41
"value name" pairs in the order READ_DATA_FILE reads them (setup.cu:17-58)."""

ORDER = [
    ("Tend", 1000), ("dx", 2e-9), ("dy", 2e-9), ("dz", 2e-9), ("num_trials", 1), ("min_trials", 0),
    ("max_trials", 1), ("TE_TM", 0), ("NCELLX", 64), ("NCELLY", 64), ("NCELLZ", 64), ("NcpmlX", 10),
    ("NcpmlY", 10), ("NcpmlZ", 10), ("inf_disp_slab", 0), ("nano_sphere", 1), ("nano_sphere_radius", 2e-8),
    ("material", 3), ("first_medium", 1), ("inc_plane", 20), ("dispersive_slab", 0), ("NUM_freq", 16),
    ("WL_or_freq", 1), ("NUM_freq_inc", 1), ("Tend_inc", 0), ("Snap_in", 0), ("Freq_start", 0),
    ("Test_offset", 0), ("Spect_loc", 0), ("BandWidth", 6.79e14), ("NONLOCAL", 0), ("Diverge_Gradient", 0),
    ("t_skip", 100), ("x_skip", 1), ("y_skip", 1), ("z_skip", 1), ("Nonlocalend", 0.0), ("MAX_AMP", 1.0),
    ("TimeFactor", 1.7320508075688772), ("lam_0", 6e-7), ("StaticField", 0),
]


def write_input(path, **over):
    vals = dict(ORDER)
    vals.update(over)
    with open(path, "w") as f:
        f.write("FDTD_input\n")
        for name, _ in ORDER:
            v = vals[name]
            f.write(("%.17g" % v if isinstance(v, float) else "%d" % v) + " " + name + "\n")
    return vals
