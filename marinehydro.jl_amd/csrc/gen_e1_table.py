"""Generates the power-series table shared by the oracle and the HIP kernels:
E1(z) = -gamma - log z - sum_{n>=1} a_n z^n,  a_n = (-1)^n / (n n!),
evaluated by Horner with degree N(ceil|z|): the first omitted term is below
2^-55 of the largest term at radius ceil|z|.  Prints C initialisers; the
including file defines E1_QUAL (storage qualifiers)."""
import mpmath as mp

mp.mp.dps = 60
NMAX = 0
deg = []
for R in range(0, 41):
    Rr = mp.mpf(max(R, 1))
    terms = [abs(Rr ** n / (n * mp.factorial(n))) for n in range(1, 400)]
    big = max(max(terms), mp.mpf(1))
    N = next(n for n in range(1, 399) if terms[n] < big * mp.mpf(2) ** -55)  # terms[n] is term n+1
    deg.append(N)
    NMAX = max(NMAX, N)
coef = [(-1) ** n / (n * mp.factorial(n)) for n in range(1, NMAX + 1)]
print(f"#define E1_SERIES_NMAX {NMAX}")
print("E1_QUAL const int E1_DEG[41] = {" + ", ".join(str(d) for d in deg) + "};")
print("E1_QUAL const double E1_A[E1_SERIES_NMAX] = {")
for i in range(0, NMAX, 3):
    print("    " + ", ".join(mp.nstr(c, 17, min_fixed=1, max_fixed=0) for c in coef[i:i + 3]) + ",")
print("};")
