/*
 * Marker rheology and density on the device: viscalc (mark3mg.c:1114-1681)
 * and the density rules of ronurecalc (dencalc1 mark3mg.c:3006-3024, the
 * phase transitions of mark3mg.c:530-641).  Included by i3g_markers.hip
 * only.  Double precision with the reference's operation order, so that a
 * marker's viscosity differs from the reference's only where the device
 * exp/pow/asin differ from the host's in the last bit.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/i3g.h"

/* viscalc's globals */
struct VisPrm {
	double nubeg, nuend, timestep, timesum, convend;
};

/* viscalc of one marker of rock R (type mm2) at T mtk (K), P mpb (bar),
 * strain-rate and stress invariants eii, sii, strain-rate correction rii
 * and dissipation dis; me, mrr, mrn are the marker's accumulated strain
 * (marke), grain roughness (markrr) and dislocation share (markrn), in/out.
 *   brittle (:1148-1199): Mohr-Coulomb strength a + b P, a and b weakened
 *     linearly between strains e0 and e1, open cracks at P < 0;
 *   ductile (:1263-1623): A constant, B exp((dh + dv P)/RT) with the two
 *     test-creep forms, C power law, D power law with exp((dh + dv P)/RT),
 *     F Karato & Wu (nu0 < 0), G the grain-size model of the mantle rocks
 *     (dh < 0: pinned-state roughness growth and damage past convend,
 *     diffusion creep r^3 / BB / 2, dislocation creep capped by the
 *     brittle strength);
 *   yield flag on the strain, dislocation share, limits (:1640-1681).
 * The Peierls branch and the melt pathway are disabled in the reference
 * (:1205, :1148 reads a buffer only the disabled melt extraction fills).
 * Where G reads the brittle strength without the brittle branch having run
 * (the reference leaves it unset) it is taken as +inf. */
__device__ inline double viscalc_dev(const i3g_rock &R, int mm2, double mtk, double mpb, double eii, double sii,
				     double rii, double dis, float &me, float &mrr, float &mrn, const VisPrm &P)
{
#pragma clang fp contract(off)
	double rt = 8.314 * mtk, epsin = eii, sigin = sii;
	double strain = fabsf(me);
	double lamb = 1.0, abrit = R.a0, bbrit = R.b0;
	if (strain > R.e1) {
		abrit = R.a1;
		bbrit = R.b1;
	} else if (strain > R.e0 && R.e1 > R.e0) {
		abrit = R.a0 + (R.a1 - R.a0) * (strain - R.e0) / (R.e1 - R.e0);
		bbrit = R.b0 + (R.b1 - R.b0) * (strain - R.e0) / (R.e1 - R.e0);
	}
	double mpf = mpb * 1e+5, nubrit = 0, sbrit = INFINITY;
	if (epsin > 0 && (bbrit > 0 || abrit * lamb > 0) && mm2 > 1) {
		sbrit = abrit + bbrit * mpf * lamb;
		if (mpf < 0) sbrit = abrit + mpf;
		if (sbrit > 0) {
			nubrit = 1.0 / (0.5 * sbrit / epsin / rii);
		} else {
			nubrit = 1.0 / R.n0;
			sbrit = 0;
		}
	}
	double nunewt = 0, nupowl = 0, e1, k1, sduct;
	if (R.nu0 != 0) {
		/* A */
		if (R.dh == 0 && R.dv == 0 && (R.ss == 0 || R.mm == 1.0)) nunewt = 1.0 / R.nu0;
		/* B */
		if ((R.dh != 0 || R.dv != 0) && (R.ss == 0 || R.mm == 1.0)) {
			e1 = (R.dh + R.dv * mpb) / rt;
			if (e1 > 150.0) e1 = 150.0;
			if (R.dh < 0 && R.dv > 0) {
				e1 = R.dh * (mtk - R.dv);
				if (e1 < -150.0) e1 = -150.0;
			}
			if (R.dh < 0 && R.dv < 0) {
				e1 = (-R.dh) * (1.0 - (mtk - (-R.dv)) / (-R.dv)) / 8.314 / (-R.dv);
				if (e1 > 150.0) e1 = 150.0;
			}
			nunewt = 1.0 / (R.nu0 * exp(e1));
		}
		/* C */
		if ((R.dh == 0 && R.dv == 0) && R.ss != 0 && R.mm != 1.0) {
			k1 = R.nu0 / pow(R.ss, R.mm - 1.0);
			nunewt = 1.0 / (0.5 * k1);
			sduct = pow(R.nu0 * epsin, 1.0 / R.mm);
			if (epsin > 0) nupowl = 1.0 / (0.5 * sduct / epsin);
		}
		/* D */
		if (R.nu0 > 0 && R.dh > 0 && (R.dh != 0 || R.dv != 0) && R.ss != 0 && R.mm != 1.0) {
			e1 = (R.dh + R.dv * mpb) / rt;
			if (e1 > 150.0) e1 = 150.0;
			e1 = exp(e1);
			k1 = R.nu0 / pow(R.ss, R.mm - 1.0);
			nunewt = 1.0 / (0.5 * k1 * e1);
			sduct = pow(R.nu0 * e1 * epsin, 1.0 / R.mm);
			if (epsin > 0) nupowl = 1.0 / (0.5 * sduct / epsin);
		}
		/* F */
		if (R.nu0 < 0) {
			double mnu2 = 1.0;
			e1 = (R.ss + mpb * R.ll) / rt;
			if (e1 > 150.0) e1 = 150.0;
			e1 = exp(e1);
			nunewt = 1.0 / (0.5 * pow(mnu2 * 1e-3 / 5e-10, 2.5) * 8e+10 / R.s1 * e1);
			nupowl = 0;
			if (epsin > 0) {
				e1 = (R.dh + R.dv * mpb) / rt / R.mm;
				if (e1 > 150.0) e1 = 150.0;
				e1 = exp(e1);
				nupowl = 1.0 / (0.5 * pow(epsin, 1.0 / R.mm - 1.0) * 8e+10 / pow(-R.nu0, 1.0 / R.mm) * e1);
			}
		}
		/* G */
		if (R.dh < 0) {
			epsin = eii;
			sigin = sii;
			double dissip = dis * (double)mrn;
			double pival = 2.0 * asin(1.0);
			double r, r0;
			r = r0 = (double)mrr;
			if (r0 <= 0) {
				mrr = (float)(1e-3 * 2.0 / pival);
				r = r0 = (double)mrr;
			}
			double dtsum = 0, drlim = 0.1, dt0 = P.timestep, dt;
			double n = R.mm;
			e1 = (-R.dh + mpb * R.dv) / rt;
			if (e1 > 150.0) e1 = 150.0;
			e1 = exp(-e1);
			double AA = R.nu0 * e1;
			double m = 3.0;
			e1 = (R.ss + mpb * R.ll) / rt;
			if (e1 > 150.0) e1 = 150.0;
			e1 = exp(-e1);
			double BB = R.s1 * e1;
			BB = BB / pow(pival / 2.0, m);
			if (P.timesum > P.convend) {
				int j = 0;
				do {
					double q = 4.0, p = 2.0, phi1 = 0.4, phi2 = 1.0 - phi1, eta = 3.0 * phi1 * phi2, gammaI = 1.0;
					double Eggrow = 300e+3, Aggrow = 2e+4 * pow(1e-6, p), Vggrow = R.ll;
					e1 = (Eggrow + mpb * Vggrow) / rt;
					if (e1 > 150.0) e1 = 150.0;
					e1 = exp(-e1);
					double GG = Aggrow * e1;
					double Gfac = 100.0;
					double GI = GG / Gfac * q / p * pow(r0, q - p);
					double f0 = 1e-3;
					e1 = 2.0 * pow(mtk / 1000, 2.9) - 2.0;
					if (e1 > 150.0) e1 = 150.0;
					e1 = exp(-e1);
					double fG = f0 * e1, fI = fG, Psi = dissip;
					(void)phi2;
					double DlnrDt = (eta * GI / q / pow(r0, q - 1.0) - fI * r0 * r0 / gammaI / eta * Psi) / r0;
					dt = dt0 - dtsum;
					if (fabs(DlnrDt * dt) > drlim) dt = dt * drlim / fabs(DlnrDt * dt);
					r = r0 * exp(DlnrDt * dt);
					if (r > 1e-2) r = 1e-2;
					if (r < 1e-6) r = 1e-6;
					j = j + 1;
					dtsum = dtsum + dt;
					r0 = r;
				} while (dtsum < dt0 && j < 50);
				mrr = (float)r;
			}
			nunewt = pow(r, m) / BB / 2.0;
			nupowl = 0;
			if (epsin > 0) {
				double epsinpowl = epsin, siginpowl = pow(epsinpowl / AA, 1.0 / n);
				nupowl = siginpowl / 2.0 / epsinpowl;
				if (sigin > sbrit) nupowl = 2.0 * AA * pow(sbrit, n - 1);
				else nupowl = 1.0 / nupowl;
			}
			nunewt = 1.0 / nunewt;
		}
	}
	double nueff = nunewt + nupowl;
	if (0 > nueff) nueff = 0;
	if (nubrit > nueff) {
		nueff = nubrit;
		if (me == 0) me = 1e-20f;
		me = fabsf(me);
	} else {
		me = -fabsf(me);
	}
	if (R.dh < 0 && P.timesum > 1e+11) mrn = (float)(nupowl / nueff);
	if (nueff <= 0) nueff = 1.0 / R.n1;
	nueff = 1.0 / nueff;
	if (nueff < R.n0) nueff = R.n0;
	if (nueff > R.n1) nueff = R.n1;
	if (nueff < P.nubeg) nueff = P.nubeg;
	if (nueff > P.nuend) nueff = P.nuend;
	return nueff;
}

/* second invariant (0.5 (a0^2 + a1^2 + a2^2) + a3^2 + a4^2 + a5^2)^(1/2) of
 * o[b..b+5], the order of mark3mg.c:69-70 */
__device__ inline double inv2_dev(const double *o, int b)
{
#pragma clang fp contract(off)
	return pow(0.5 * (o[b] * o[b] + o[b + 1] * o[b + 1] + o[b + 2] * o[b + 2]) + o[b + 3] * o[b + 3] +
			   o[b + 4] * o[b + 4] + o[b + 5] * o[b + 5],
		   0.5);
}

/* dencalc1 (mark3mg.c:3006-3024); bb = its adiabatic term eps1[20] */
__device__ inline double dencalc_dev(const i3g_rock &r, int densimod, double mtk, double mpb, double &bb)
{
#pragma clang fp contract(off)
	bb = 0;
	if (densimod == 0) return r.ro;
	bb = r.bro / (1.0 - r.bro * (mtk - 298.15));
	return r.ro * (1.0 - r.bro * (mtk - 298.15)) * (1.0 + r.aro * (mpb - 1.0) * 1e-3);
}

/* density factor of ronurecalc's phase transitions (mark3mg.c:530-641):
 * types 7-8 and 16-18 basalt -> garnet granulite -> eclogite (Ito and
 * Kennedy), coesite -> stishovite, perovskite; types 9-14 olivine ->
 * spinel (Katsura and Ito), perovskite; the factor is damped linearly
 * between teclmin and teclmax and 1 below teclmin */
__device__ inline double phase_rokf_dev(int mm2, double mtk, double mpb, double teclmin, double teclmax)
{
#pragma clang fp contract(off)
	bool ecl = (mm2 >= 7 && mm2 <= 8) || (mm2 >= 16 && mm2 <= 18), man = mm2 >= 9 && mm2 <= 14;
	if (!ecl && !man) return 1.0;
	double rokf = 1.0;
	if (ecl) {
		double p_ga_in = -9222.0 + mtk * 14.0;
		if (p_ga_in < 2000.0) p_ga_in = 2000.0;
		double p_pl_out = -1460.0 + mtk * 20.0;
		if (p_pl_out < 12000.0) p_pl_out = 12000.0;
		if (mpb > p_ga_in) {
			if (mpb >= p_pl_out) rokf *= 1.16;
			else rokf *= 1.0 + 0.16 * (mpb - p_ga_in) / (p_pl_out - p_ga_in);
		}
		double p_st_in = 59100.0 + mtk * 22.6;
		if (mpb > p_st_in) rokf *= 1.06;
		double p_sp_out = 354000.0 - mtk * 40.0, p_pv_in = 352000.0 - mtk * 40.0;
		if (mpb > p_pv_in) {
			if (mpb >= p_sp_out) rokf *= 1.08;
			else rokf *= 1.0 + 0.08 * (mpb - p_pv_in) / (p_sp_out - p_pv_in);
		}
	} else {
		double p_ol_out = 91000.0 + mtk * 27.0, p_sp_in = 66000.0 + mtk * 39.0;
		if (p_sp_in > p_ol_out - 2000.0) p_sp_in = p_ol_out - 2000.0;
		if (mpb > p_sp_in) {
			if (mpb >= p_ol_out) rokf *= 1.06;
			else rokf *= 1.0 + 0.06 * (mpb - p_sp_in) / (p_ol_out - p_sp_in);
		}
		double p_sp_out = 304000.0 - mtk * 40.0, p_pv_in = 302000.0 - mtk * 40.0;
		if (mpb > p_pv_in) {
			if (mpb >= p_sp_out) rokf *= 1.11;
			else rokf *= 1.0 + 0.11 * (mpb - p_pv_in) / (p_sp_out - p_pv_in);
		}
	}
	if (mtk < teclmax) {
		if (mtk < teclmin) rokf = 1.00;
		else rokf = 1.0 + (rokf - 1.0) * (mtk - teclmin) / (teclmax - teclmin);
	}
	return rokf;
}

/* ---- partially molten rocks (meltpart11, meltpart0) -------------------- */

/* solidus depression by water (Katz et al. 2003, mark3mg.c:1950-1990):
 * dT = 43 X^0.75, X = mwa / (0.01 + F (1 - 0.01)) up to the saturation
 * 12 P^0.6 + P (P in GPa) */
__device__ inline double melt_wet_dt(double mwa, double F, double pg)
{
#pragma clang fp contract(off)
	double sat = 12.0 * pow(pg, 0.6) + 1.0 * pg;
	double X = mwa / (0.01 + F * (1.0 - 0.01));
	return 43.0 * pow(X > sat ? sat : X, 0.75);
}

/* rock family of the melting model (mark3mg.c:1889-1953): 0 granite
 * (sediments, upper crust), 1 basalt, 2 gabbro, 3 peridotite, -1 none */
__device__ inline int melt_family(int mm2)
{
	switch (mm2) {
	case 3: case 4: case 5: case 15: case 17: case 23: case 24: case 25: case 35: case 37: return 0;
	case 7: case 16: case 18: case 27: case 36: case 38: return 1;
	case 6: case 8: case 26: case 28: return 2;
	case 9: case 10: case 11: case 12: case 14: case 29: case 30: case 31: case 32: case 34: return 3;
	default: return -1;
	}
}

/* meltpart11 (mark3mg.c:1872-2070): melt fraction F and latent heat of
 * rock type mm2 with water content mwa at mtk (K), mpb (bar); maxpmelt (Pa)
 * stops dry peridotite melting.  Peridotite follows Katz et al. (2003),
 * the water's solidus depression found by bisection over F (at most 100
 * halvings, to 1e-3 K); F is finally linear between the solidus and the
 * (effective) liquidus. */
__device__ inline void meltpart11_dev(int mm2, double mtk, double mpb, double mwa, double maxpmelt, double &F,
				      double &H)
{
#pragma clang fp contract(off)
	double ykm = mpb * 3e-3, ts = 0, tl = 0, hl = 0;
	F = 0;
	H = 0;
	int fam = ykm > 0 ? melt_family(mm2) : -1;
	if (fam == 0) {
		double a = ykm + 1.609;
		ts = ykm < 36.0 ? 889.0 + 536.6 / a + 18.21 / a / a : 831.3 + 2.0 * ykm;
		tl = 1262.0 + 3.0 * ykm;
		hl = 300000.0;
	} else if (fam == 1) {
		double a = ykm + 10.63;
		ts = ykm < 48.0 ? 972.6 - 2111.0 / a + 70033.0 / a / a : 935.4 + 0.1162 * ykm + 0.006937 * ykm * ykm;
		tl = 1423.15 + 3.5 * ykm;
		hl = 380000.0;
	} else if (fam == 2) {
		ts = 1327.15 + 3.02 * ykm;
		tl = 1423.15 + 3.5 * ykm;
		hl = 380000.0;
	} else if (fam == 3 && (mwa > 0 || !(mpb * 1e+5 > maxpmelt))) {
		double pg = mpb * 1e-4, dlo = 0, dhi = 0;
		int steps = 1;
		if (mwa > 0) {
			dlo = dhi = melt_wet_dt(mwa, 0.0, pg);
			double t0 = 273.15 + 1085.7 + 132.9 * pg - 5.1 * pg * pg - dlo;
			if (mtk >= t0) {
				dlo = melt_wet_dt(mwa, 1.0, pg);
				steps = 100;
			}
		}
		for (int it = 0; it < steps; it++) {
			double dt = (dlo + dhi) / 2.0, f = 0;
			ts = 273.15 + 1085.7 + 132.9 * pg - 5.1 * pg * pg - dt;
			double tlz = 273.15 + 1475.0 + 80.0 * pg - 3.2 * pg * pg - dt;
			tl = 273.15 + 1780.0 + 45.0 * pg - 2.0 * pg * pg - dt;
			if (mtk >= ts) {
				if (mtk >= tl) {
					f = 1.0;
				} else {
					double fc = 0.15 / (0.5 + 0.08 * pg);
					double tc = pow(fc, 1.0 / 1.5) * (tlz - ts) + ts;
					if (mtk > tc) f = fc + (1.0 - fc) * pow((mtk - tc) / (tl - tc), 1.5);
					else f = pow((mtk - ts) / (tlz - ts), 1.5);
					tl = (mtk - ts) / f + ts;
				}
			}
			if (mwa > 0) {
				if (melt_wet_dt(mwa, f, pg) > dt) dlo = dt;
				else dhi = dt;
			}
			if (fabs(dhi - dlo) < 0.001) break;
		}
		hl = 400000.0;
	}
	if (tl != 0) {
		double f = (mtk - ts) / (tl - ts);
		f = f < 0 ? 0 : (f > 1.0 ? 1.0 : f);
		F = f;
		H = hl * f;
	}
}

/* meltpart0 (mark3mg.c:2152-2340) for a marker of melt type mm2 (> 20)
 * whose solid is S (type mm2 - 20, 11 for 34): the expansion term bb,
 * density ro, heat capacity cp and conductivity kt of the rock mixed by
 * the melt fraction less the extracted melt mex; densimod 1 with the
 * solid's eclogite (27-28, 36-38) or mantle (29-34) transition factor;
 * inside the melting interval the latent heat's central differences in P
 * (300 bar) and T (10 K) are added to bb (within +-1e-2) and cp
 * (100..5e4).  xm is the melt fraction of the last evaluation (the
 * reference's eps1[21]).  mm2 = 21: every output zero. */
__device__ inline void meltpart0_dev(const i3g_rock &A, const i3g_rock &S, int mm2, double mtk, double mpb, double mwa,
				     double mex, int densimod, double teclmin, double teclmax, double maxpmelt,
				     double &bb, double &ro, double &cp, double &kt, double &xm)
{
#pragma clang fp contract(off)
	bb = ro = cp = kt = xm = 0;
	if (mm2 <= 21) return;
	double H;
	meltpart11_dev(mm2, mtk, mpb, mwa, maxpmelt, xm, H);
	double f = xm - mex;
	if (f < 0) f = 0;
	bb = (A.bro * f + S.bro * (1.0 - f)) / (1.0 - (A.bro * f + S.bro * (1.0 - f)) * (mtk - 298.15));
	if (densimod == 0) {
		ro = A.ro * f + S.ro * (1.0 - f);
	} else {
		int fam = (mm2 >= 27 && mm2 <= 28) || (mm2 >= 36 && mm2 <= 38) ? 7 : (mm2 >= 29 && mm2 <= 34 ? 9 : 0);
		double rokf = fam ? phase_rokf_dev(fam, mtk, mpb, teclmin, teclmax) : 1.0;
		ro = f * A.ro * (1.0 - A.bro * (mtk - 298.15)) * (1.0 + A.aro * (mpb - 1.0) * 1e-3) +
		     (1.0 - f) * rokf * S.ro * (1.0 - S.bro * (mtk - 298.15)) * (1.0 + S.aro * (mpb - 1.0) * 1e-3);
	}
	cp = A.cp * f + S.cp * (1.0 - f);
	kt = ((A.kt + A.kf / (mtk + 77.0)) * exp(A.kp * mpb)) * f + ((S.kt + S.kf / (mtk + 77.0)) * exp(S.kp * mpb)) * (1.0 - f);
	if (f > 0 && f < 1.0) {
		double Hm, Hp, x;
		meltpart11_dev(mm2, mtk, mpb - 300.0, mwa, maxpmelt, x, Hm);
		meltpart11_dev(mm2, mtk, mpb + 300.0, mwa, maxpmelt, x, Hp);
		bb += (Hm - Hp) * (ro / (mtk * 2.0 * 300.0 * 1e+5));
		bb = bb < -1e-2 ? -1e-2 : (bb > 1e-2 ? 1e-2 : bb);
		meltpart11_dev(mm2, mtk + 10.0, mpb, mwa, maxpmelt, x, Hp);
		meltpart11_dev(mm2, mtk - 10.0, mpb, mwa, maxpmelt, xm, Hm);
		cp += (Hp - Hm) / (2.0 * 10.0);
		cp = cp < 1e+2 ? 1e+2 : (cp > 5e+4 ? 5e+4 : cp);
	}
}
