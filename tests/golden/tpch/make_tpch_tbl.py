"""Writes the small TPC-H-format fixture of this directory (region, nation, supplier, customer,
part, partsupp, orders, lineitem .tbl files): the column layout and key relationships of the
TPC-H schema (every foreign key refers to an existing row; four suppliers per part), at a tiny
scale, with seeded values. Only the integer fields are read by the tests (tbl2db parses the
leading integer of a field); text fields hold placeholders. Run from any directory:

    python tests/golden/tpch/make_tpch_tbl.py
"""
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))
N_REGION, N_NATION, N_SUPP, N_CUST, N_PART, N_ORDERS = 5, 25, 20, 60, 80, 300


def write(name, rows):
    with open(os.path.join(HERE, name + ".tbl"), "w") as f:
        for r in rows:
            f.write("|".join(str(x) for x in r) + "|\n")


def main():
    rng = random.Random(20240808)
    write("region", [(r, f"REGION{r}", "c") for r in range(N_REGION)])
    write("nation", [(n, f"NATION{n}", n % N_REGION, "c") for n in range(N_NATION)])
    write("supplier", [(s, f"Supplier#{s}", "addr", rng.randrange(N_NATION), "11-111",
                        f"{rng.uniform(0, 9999):.2f}", "c") for s in range(1, N_SUPP + 1)])
    write("customer", [(c, f"Customer#{c}", "addr", rng.randrange(N_NATION), "22-222",
                        f"{rng.uniform(0, 9999):.2f}", "BUILDING", "c")
                       for c in range(1, N_CUST + 1)])
    write("part", [(p, "part", "Manufacturer#1", "Brand#11", "ECONOMY ANODIZED STEEL",
                    rng.randint(1, 50), "SM BOX", f"{900 + p / 10:.2f}", "c")
                   for p in range(1, N_PART + 1)])
    ps = []
    for p in range(1, N_PART + 1):  # four distinct suppliers per part, as TPC-H
        for s in rng.sample(range(1, N_SUPP + 1), 4):
            ps.append((p, s, rng.randint(1, 9999), f"{rng.uniform(1, 1000):.2f}", "c"))
    write("partsupp", ps)
    orders, lines = [], []
    for o in range(1, N_ORDERS + 1):
        okey = o * 4 - 3  # TPC-H order keys are sparse
        orders.append((okey, rng.randint(1, N_CUST), "O", f"{rng.uniform(1, 9e5):.2f}",
                       "1995-03-15", "1-URGENT", "Clerk#1", 0, "c"))
        for ln in range(1, rng.randint(1, 7) + 1):
            p = rng.randint(1, N_PART)
            s = rng.choice([x[1] for x in ps[(p - 1) * 4:(p - 1) * 4 + 4]])
            lines.append((okey, p, s, ln, rng.randint(1, 50), f"{rng.uniform(900, 9e4):.2f}",
                          "0.05", "0.02", "N", "O", "1995-04-01", "1995-04-02", "1995-04-03",
                          "NONE", "AIR", "c"))
    write("orders", orders)
    write("lineitem", lines)


if __name__ == "__main__":
    main()
