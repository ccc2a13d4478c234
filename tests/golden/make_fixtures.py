"""Extract the reference's golden data into small fixtures (run where the
reference tree is mounted; the fixtures are committed so no test reads it).

- f1_tdvp_ising.csv: tdvp_benchmark_results.csv Ising rows (exact, 1TDVP,
  2TDVP <Z> at sites 1, 3, 6; L=6, dt=0.05, T=2), produced by
  test/benchmark_tdvp.jl:94-216.
- f3_lindblad_exact.csv: the exact (QuTiP mesolve) <Z_i> columns of
  00_safety_checks_full_algorithms/analogTJM_check/non_long_range_check/
  results/tjm_vs_qutip_data.csv (L=6 Ising J=1 h=0.5, raising gamma=0.1,
  dt=0.05, T=2; written by benchmark_analog_tjm.jl:90-227), plus its TJM
  columns (500 trajectories) for reference.
- f2_large_ising.csv: large_ising_results.csv (L=30 Ising J=1 g=0.5, chi=32,
  noise-free 2TDVP, test/simulate_large_ising.jl:31-48).
"""
import csv
import os

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    rows = list(csv.DictReader(open(os.path.join(REF, "tdvp_benchmark_results.csv"))))
    with open(os.path.join(OUT, "f1_tdvp_ising.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Model", "Method", "Time", "Site", "ExpVal"])
        for r in rows:
            if r["Model"].startswith("Ising"):
                w.writerow([r["Model"], r["Method"], r["Time"], r["Site"], r["ExpVal"]])
    src = os.path.join(REF, "00_safety_checks_full_algorithms/analogTJM_check/non_long_range_check/results/"
                            "tjm_vs_qutip_data.csv")
    with open(src) as f, open(os.path.join(OUT, "f3_tjm_vs_lindblad.csv"), "w", newline="") as g:
        w = csv.writer(g)
        w.writerow(["time"] + [f"tjm_z{i}" for i in range(1, 7)] + ["tjm_stag"] +
                   [f"exact_z{i}" for i in range(1, 7)] + ["exact_stag"])
        for line in f:
            line = line.strip()
            if line:
                w.writerow(line.split(","))
    with open(os.path.join(REF, "large_ising_results.csv")) as f, \
            open(os.path.join(OUT, "f2_large_ising.csv"), "w") as g:
        g.write(f.read())


if __name__ == "__main__":
    main()
