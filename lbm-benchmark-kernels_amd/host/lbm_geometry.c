/*
 * lbm_geometry.c -- geometry masks: the reference's five generators
 * (src/Geometry.c:37-281) and a seeded porous medium of solid spheres.
 */
#include "lbm_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

static const char * g_geoNames[] = { "box", "channel", "pipe", "blocks", "fluid", "porous" };

static void geo_error(const char * msg)
{
	printf("ERROR: %s\n", msg);
	exit(1);
}

static inline int at(const int d[3], int x, int y, int z) { return (z * d[1] + y) * d[0] + x; }

static void count_cells(LatticeDesc * ld)
{
	const int n = ld->Dims[0] * ld->Dims[1] * ld->Dims[2];
	ld->nObst = ld->nFluid = ld->nInlet = ld->nOutlet = 0;
	for (int i = 0; i < n; ++i) {
		switch (ld->Lattice[i]) {
			case LBM_CELL_OBSTACLE: ld->nObst++; break;
			case LBM_CELL_FLUID:    ld->nFluid++; break;
			case LBM_CELL_INLET:    ld->nInlet++;  ld->nFluid++; break;
			case LBM_CELL_OUTLET:   ld->nOutlet++; ld->nFluid++; break;
			default: geo_error("invalid cell type in lattice.");
		}
	}
}

/* SplitMix64 (Steele, Lea, Flood 2014): state += golden gamma, then two
 * xor-shift-multiply rounds. */
static uint64_t splitmix64(uint64_t * s)
{
	uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

static double uniform01(uint64_t * s)
{
	return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);   /* 53 bits -> [0,1) */
}

static void place_spheres(LatticeDesc * ld, GeoPorous * gp)
{
	const int * d = ld->Dims;
	const int X = d[0], Y = d[1], Z = d[2];
	const double r = gp->Radius, r2 = r * r;
	const double total = (double)X * Y * Z;
	uint64_t state = gp->Seed;
	long fluid = ld->nFluid;
	const long maxSpheres = 100000000L;
	gp->nSpheres = 0;
	if (!(r > 0.0)) geo_error("porous geometry: sphere radius must be > 0.");
	if (!(gp->Porosity >= 0.0 && gp->Porosity < 1.0)) geo_error("porous geometry: porosity must be in [0, 1).");
	while (fluid / total > gp->Porosity) {
		if (gp->nSpheres >= maxSpheres) geo_error("porous geometry: porosity not reached.");
		const double cx = uniform01(&state) * X;
		const double cy = uniform01(&state) * Y;
		const double cz = uniform01(&state) * Z;
		const int z0 = (int)ceil(cz - r), z1 = (int)floor(cz + r);
		const int y0 = (int)ceil(cy - r), y1 = (int)floor(cy + r);
		const int x0 = (int)ceil(cx - r), x1 = (int)floor(cx + r);
		for (int z = z0 < 0 ? 0 : z0; z <= z1 && z < Z; ++z)
		for (int y = y0 < 0 ? 0 : y0; y <= y1 && y < Y; ++y)
		for (int xi = x0; xi <= x1; ++xi) {
			const int x = ((xi % X) + X) % X;               /* periodic in x */
			const double dx = xi - cx, dy = y - cy, dz = z - cz;
			if (dx * dx + dy * dy + dz * dz > r2) continue;
			int * c = &ld->Lattice[at(d, x, y, z)];
			if (*c != LBM_CELL_OBSTACLE) { *c = LBM_CELL_OBSTACLE; --fluid; }
		}
		gp->nSpheres++;
	}
}

void GeoCreateByType(GEO_TYPES type, void * typeDetails, int dims[3], int periodic[3], LatticeDesc * ld)
{
	if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) geo_error("dimensions must be > 0.");
	if (type < GEO_TYPE_MIN || type > GEO_TYPE_MAX) geo_error("unknown geometry specified.");
	const int X = dims[0], Y = dims[1], Z = dims[2];

	ld->Dims[0] = X; ld->Dims[1] = Y; ld->Dims[2] = Z;
	ld->nCells = X * Y * Z;
	/* the stored flags are the requested ones, taken before the generator
	 * forces fluid faces below (src/Geometry.c:115-139) */
	ld->PeriodicX = periodic[0];
	ld->PeriodicY = periodic[1];
	ld->PeriodicZ = periodic[2];
	ld->Name = g_geoNames[type];
	ld->Lattice = (LatticeT *)malloc(sizeof(LatticeT) * (size_t)X * Y * Z);
	if (!ld->Lattice) geo_error("allocating the geometry mask failed.");
	for (int i = 0; i < X * Y * Z; ++i) ld->Lattice[i] = LBM_CELL_FLUID;

	if (type == GEO_TYPE_CHANNEL || type == GEO_TYPE_BLOCKS || type == GEO_TYPE_PIPE || type == GEO_TYPE_POROUS)
		periodic[0] = 1;
	else if (type == GEO_TYPE_FLUID)
		periodic[0] = periodic[1] = periodic[2] = 1;
	if (type == GEO_TYPE_POROUS) ld->PeriodicX = 1;

	/* faces: fluid where periodic, wall otherwise; x faces first, then y,
	 * then z, so edges follow the last face written (src/Geometry.c:141-178) */
	for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) {
		ld->Lattice[at(dims, 0, y, z)]     = periodic[0] ? LBM_CELL_FLUID : LBM_CELL_OBSTACLE;
		ld->Lattice[at(dims, X - 1, y, z)] = periodic[0] ? LBM_CELL_FLUID : LBM_CELL_OBSTACLE;
	}
	for (int z = 0; z < Z; ++z) for (int x = 0; x < X; ++x) {
		ld->Lattice[at(dims, x, 0, z)]     = periodic[1] ? LBM_CELL_FLUID : LBM_CELL_OBSTACLE;
		ld->Lattice[at(dims, x, Y - 1, z)] = periodic[1] ? LBM_CELL_FLUID : LBM_CELL_OBSTACLE;
	}
	for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
		ld->Lattice[at(dims, x, y, 0)]     = periodic[2] ? LBM_CELL_FLUID : LBM_CELL_OBSTACLE;
		ld->Lattice[at(dims, x, y, Z - 1)] = periodic[2] ? LBM_CELL_FLUID : LBM_CELL_OBSTACLE;
	}

	if (type == GEO_TYPE_PIPE) {
		/* circular cross section in y-z (src/Geometry.c:183-200) */
		const double cz = Z / 2.0 - 0.5, cy = Y / 2.0 - 0.5;
		const double dmin = Y < Z ? Y : Z;
		const double r2 = (dmin / 2 - 1) * (dmin / 2 - 1);
		for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) {
			if ((z - cz) * (z - cz) + (y - cy) * (y - cy) < r2) continue;
			for (int x = 0; x < X; ++x) ld->Lattice[at(dims, x, y, z)] = LBM_CELL_OBSTACLE;
		}
	}
	else if (type == GEO_TYPE_BLOCKS) {
		/* solid cubes of edge b on a 2b pitch (src/Geometry.c:201-241) */
		const int b = typeDetails ? *(int *)typeDetails : 8;
		if (b > 0) {
			int dmin = X < Y ? X : Y;
			if (Z < dmin) dmin = Z;
			if (b > dmin / 2) geo_error("block size for geometry must be > 0 and smaller than half of the smalest dimension.");
			for (int z = 0; z < Z; ++z) {
				if (z % (2 * b) < b) continue;
				for (int y = 0; y < Y; ++y) {
					if (y % (2 * b) < b) continue;
					for (int x = 0; x < X; ++x)
						if (x % (2 * b) >= b) ld->Lattice[at(dims, x, y, z)] = LBM_CELL_OBSTACLE;
				}
			}
		}
	}

	count_cells(ld);

	if (type == GEO_TYPE_POROUS) {
		GeoPorous defaults = { 6.0, 0.8, 42, 0 };
		GeoPorous * gp = typeDetails ? (GeoPorous *)typeDetails : &defaults;
		place_spheres(ld, gp);
		count_cells(ld);
	}
}

void GeoCreateByStr(const char * s, int dims[3], int periodic[3], LatticeDesc * ld)
{
	if (!strncasecmp("channel", s, 7)) {
		GeoCreateByType(GEO_TYPE_CHANNEL, NULL, dims, periodic, ld);
	}
	else if (!strncasecmp("box", s, 3)) {
		GeoCreateByType(GEO_TYPE_BOX, NULL, dims, periodic, ld);
	}
	else if (!strncasecmp("pipe", s, 4)) {
		GeoCreateByType(GEO_TYPE_PIPE, NULL, dims, periodic, ld);
	}
	else if (!strncasecmp("blocks", s, 6)) {
		int b = 8;
		if (strlen(s) > 7) {
			b = atoi(s + 7);
			int dmin = dims[0] < dims[1] ? dims[0] : dims[1];
			if (dims[2] < dmin) dmin = dims[2];
			if (b < 0 || b > dmin / 2)
				geo_error("block size for geometry must be > 0 and smaller than half of the smalest dimension.");
		}
		GeoCreateByType(GEO_TYPE_BLOCKS, &b, dims, periodic, ld);
	}
	else if (!strncasecmp("fluid", s, 5)) {
		GeoCreateByType(GEO_TYPE_FLUID, NULL, dims, periodic, ld);
	}
	else if (!strncasecmp("porous", s, 6)) {
		GeoPorous gp = { 6.0, 0.8, 42, 0 };
		const char * t = s + 6;
		if (*t == '-') {
			char * end;
			gp.Radius = strtod(t + 1, &end); t = end;
			if (*t == '-') { gp.Porosity = strtod(t + 1, &end); t = end; }
			if (*t == '-') { gp.Seed = strtoull(t + 1, &end, 10); t = end; }
			if (*t != '\0') geo_error("porous geometry string: porous[-<radius>[-<porosity>[-<seed>]]].");
		}
		else if (*t != '\0') geo_error("unknown geometry specified.");
		GeoCreateByType(GEO_TYPE_POROUS, &gp, dims, periodic, ld);
		printf("# porous medium: radius %.3f porosity %.6f (target %.3f) seed %llu spheres %d\n",
		       gp.Radius, (double)ld->nFluid / ld->nCells, gp.Porosity, (unsigned long long)gp.Seed, gp.nSpheres);
	}
	else {
		geo_error("unknown geometry specified.");
	}
}

void GeoFree(LatticeDesc * ld)
{
	free(ld->Lattice);
	ld->Lattice = NULL;
}
