"""Copies the input files of two of CHAMP's own CI cases from the reference
tree into fixtures (large numerical grids and orbital files gzipped), with
the energies the CI scripts compare against:

  H2      tests/CI_test/VMC-H2 (revised_vmc.inp; energy from
          VMC_H2_STO-CVB1_energy_hydrogen.sh:8-9 and :19-20)
  C4H6    tests/CI_test/VMC-C4H6-ci44_pVQZ_20_dets_12_csf_noopt
          (vmc_noopt_ci44.inp; energies from
          VMC_C4H6_ci44_pVQZ_20-dets_12-csf_energy-only_butadiene.sh:8-9, :19-20;
          the CHAMP-v2 orbital file and basis info its conversion step reads)
  C4H6_dmc  tests/CI_test/DMC-C4H6-ci1010_pVTZ-500-dets (dmc.inp; energies from
          DMC_C4H6_ci1010_pVTZ_500-dets_butadiene.sh:8-9, :16-17)
  C4H6_5k   tests/CI_test/VMC-C4H6-ci1010_pVTZ-05000-dets (vmc_optimization_5000.inp,
          one sr_n iteration over the Jastrow, determinant and orbital
          parameters; energies of the closing VMC run from
          VMC_C4H6_ci1010_pVTZ_5000-dets_butadiene_optimization.sh:8-9, :18-19)
  C4H6_500  tests/CI_test/VMC-C4H6-ci1010_pVTZ-00500-dets (the same with 500
          determinants; VMC_C4H6_ci1010_pVTZ_500-dets_butadiene_optimization.sh)
  C4H6_15k  tests/CI_test/VMC-C4H6-ci1010_pVTZ-15000-dets (the same with 15000
          determinants; VMC_C4H6_ci1010_pVTZ_15000-dets_butadiene_optimization.sh)

Run in the build container only (the reference tree is not on GPU hosts):
    python tests/golden/champ_inputs/make_fixtures.py
"""
import gzip
import json
import os
import re
import shutil

REF = "/root/reference/tests/CI_test"
HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "H2": ("VMC-H2", "VMC_H2_STO-CVB1_energy_hydrogen.sh",
           ["revised_vmc.inp", "rhf.det", "champ_v3_trexio_order_rhf.sto_cvb1.lcao", "jastrow.0", "jastrow.der",
            "pool/champ_v3_h2.xyz", "pool/champ_v3_sto_cvb1_basis_pointers.bfinfo", "pool/sto_cvb1.basis.H"]),
    "C4H6_ci44": ("VMC-C4H6-ci44_pVQZ_20_dets_12_csf_noopt",
                  "VMC_C4H6_ci44_pVQZ_20-dets_12-csf_energy-only_butadiene.sh",
                  ["vmc_noopt_ci44.inp", "champ_v2_gamess_ci44_pVQZ_cart.det", "jastrow_good_b3lyp.0", "jastrow.der",
                   "pool/champ_v3_butadiene_champ_v2_gamess_ci44_pVQZ_1_geom.xyz",
                   "pool/champ_v3_champ_v2_gamess_ci44_pVQZ_basis_pointers.bfinfo",
                   "pool/BFD-Q.basis.C", "pool/BFD-Q.basis.H", "pool/BFD.gauss_ecp.dat.C",
                   "pool/BFD.gauss_ecp.dat.H", "conversion/champ_v2_gamess_ci44_pVQZ_cart.lcao",
                   "conversion/champ_v2_gamess_ci44_pVQZ.bfinfo"]),
    "C4H6_dmc": ("DMC-C4H6-ci1010_pVTZ-500-dets", "DMC_C4H6_ci1010_pVTZ_500-dets_butadiene.sh",
                 ["dmc.inp", "TZ_1M_500.det", "champ_v3_trexio_order_ci1010_pVTZ_1_orbitals.lcao",
                  "jastrow_good_b3lyp.0", "pool/champ_v3_butadiene.xyz", "pool/champ_v3_BFD-T_basis_pointers.bfinfo",
                  "pool/BFD-T.basis.C", "pool/BFD-T.basis.H", "pool/BFD.gauss_ecp.dat.C", "pool/BFD.gauss_ecp.dat.H"]),
    "C4H6_5k": ("VMC-C4H6-ci1010_pVTZ-05000-dets", "VMC_C4H6_ci1010_pVTZ_5000-dets_butadiene_optimization.sh",
                ["vmc_optimization_5000.inp", "TZ_1M_5k.det", "champ_v3_trexio_order_ci1010_pVTZ_1_orbitals.lcao",
                 "ci1010_pVTZ_1_symmetry.sym", "jastrow_good_b3lyp.0", "jastrow.der", "pool/champ_v3_butadiene.xyz",
                 "pool/champ_v3_BFD-T_basis_pointers.bfinfo", "pool/BFD-T.basis.C", "pool/BFD-T.basis.H",
                 "pool/BFD.gauss_ecp.dat.C", "pool/BFD.gauss_ecp.dat.H"]),
    "C4H6_500": ("VMC-C4H6-ci1010_pVTZ-00500-dets", "VMC_C4H6_ci1010_pVTZ_500-dets_butadiene_optimization.sh",
                 ["vmc_optimization_500.inp", "TZ_1M_500.det", "champ_v3_trexio_order_ci1010_pVTZ_1_orbitals.lcao",
                  "ci1010_pVTZ_1_symmetry.sym", "jastrow_good_b3lyp.0", "jastrow.der", "pool/champ_v3_butadiene.xyz",
                  "pool/champ_v3_BFD-T_basis_pointers.bfinfo", "pool/BFD-T.basis.C", "pool/BFD-T.basis.H",
                  "pool/BFD.gauss_ecp.dat.C", "pool/BFD.gauss_ecp.dat.H"]),
    "C4H6_15k": ("VMC-C4H6-ci1010_pVTZ-15000-dets", "VMC_C4H6_ci1010_pVTZ_15000-dets_butadiene_optimization.sh",
                 ["vmc_optimization_15000.inp", "TZ_1M_15k.det", "champ_v3_trexio_order_ci1010_pVTZ_1_orbitals.lcao",
                  "ci1010_pVTZ_1_symmetry.sym", "jastrow_good_b3lyp.0", "jastrow.der", "pool/champ_v3_butadiene.xyz",
                  "pool/champ_v3_BFD-T_basis_pointers.bfinfo", "pool/BFD-T.basis.C", "pool/BFD-T.basis.H",
                  "pool/BFD.gauss_ecp.dat.C", "pool/BFD.gauss_ecp.dat.H"]),
}


def energies(script):
    """(cores, energy, error) of every comparison in a CI script."""
    txt = open(script).read()
    out = []
    for m in re.finditer(r"N=(\d+)\s+ReferenceEnergy=(-?[\d.]+)\s+ReferenceError=([\d.]+)", txt):
        out.append({"cores": int(m.group(1)), "energy": float(m.group(2)), "error": float(m.group(3))})
    return out


def main():
    for name, (d, script, files) in CASES.items():
        src, dst = os.path.join(REF, d), os.path.join(HERE, name)
        for f in files:
            os.makedirs(os.path.dirname(os.path.join(dst, f)), exist_ok=True)
            if os.path.getsize(os.path.join(src, f)) > 100_000:
                out = os.path.join(dst, f) + ".gz"
                if os.path.exists(out):  # gzip output carries a timestamp: keep committed fixtures as they are
                    continue
                with open(os.path.join(src, f), "rb") as fi, open(out, "wb") as raw, \
                        gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as fo:
                    shutil.copyfileobj(fi, fo)
            else:
                shutil.copy(os.path.join(src, f), os.path.join(dst, f))
        with open(os.path.join(dst, "reference_energies.json"), "w") as fh:
            json.dump({"source": f"tests/CI_test/{d}/{script}", "total_E": energies(os.path.join(src, script))},
                      fh, indent=1)
        print(name, "->", dst)


if __name__ == "__main__":
    main()
