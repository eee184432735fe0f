/*
  Cross-check driver linked against objects compiled from the reference's
  own sources (oracle/Makefile.ref).  It is test infrastructure only: it
  writes golden vectors that pin the CPU oracle and the HIP kernels.

  modes
    ref_kgen geom  <geom_file>
        one line per patch: x y z cos_alpha sin_alpha l w
        normal[3] t_strike[3] t_dip[3] (the reference's own preprocessing,
        read_geometry src/read_geometry.c:23 -> calc_quad_base_vecs
        src/geometry/geometry.c:128)
    ref_kgen kmat  <geom_file> <full_space 0|1> <out.bin>
        for every (receiver i, source j): traction (strike, dip, normal)
        at receiver i for unit strike slip and unit dip slip on source j,
        in internal units (shear modulus 1e4, nu 0.25, no MatScale), as
        GenKEntries_petsc does it (src/interact_matrix_assembly.c:130-145):
        eval_green_at_receiver + resolve_stress_on_fault, singular -> 0.
        out.bin holds double[P][P][2][3] followed by int32 iret[P][P][2]
    ref_kgen kone  <geom_file> <full_space 0|1>
        prints |K 1|_2 for the strike/strike operator at scale 1, the
        quantity of src/compress_interaction_matrix.c:954-963
    ref_kgen dc3d
        reads "alpha x y z depth dip al1 al2 aw1 aw2 d1 d2 d3 ifull"
        lines from stdin and prints the 12 DC3D outputs and iret
        (src/green/dc3d.F:576)
*/
#include "interact.h"
#include "properties.h"

extern void dc3d(double*,double*,double*,double*,double*,double*,double*,double*,
		 double*,double*,double*,double*,double*,double*,double*,double*,
		 double*,double*,double*,double*,double*,double*,double*,double*,
		 double*,int*,int*);

static void traction_pair(struct med *medium, struct flt *fault, int i, int j,
			  int full_space, double out[2][3], int iret_out[2])
{
  double slip[3], disp[3], stress[3][3];
  int s, iret;
  for (s = 0; s < 2; s++) {
    slip[0] = slip[1] = slip[2] = 0.0;
    slip[s] = 1.0;		/* unit slip, get_right_slip semantics */
    eval_green_at_receiver(fault, i, j, slip, disp, stress, &iret,
			   GC_STRESS_ONLY, FALSE, (my_boolean)full_space,
			   medium->elastic);
    iret_out[s] = iret;
    if (iret != 0) {
      out[s][0] = out[s][1] = out[s][2] = 0.0;
    } else {
      out[s][0] = resolve_stress_on_fault(stress, fault + i, STRIKE);
      out[s][1] = resolve_stress_on_fault(stress, fault + i, DIP);
      out[s][2] = resolve_stress_on_fault(stress, fault + i, NORMAL);
    }
  }
}

static void load(char *fname, int full_space, struct med **medium, struct flt **fault)
{
  *medium = (struct med *)calloc(1, sizeof(struct med));
  (*medium)->full_space = (my_boolean)full_space;
  read_geometry(fname, medium, fault, FALSE, FALSE, FALSE);
}

int main(int argc, char **argv)
{
  struct med *medium;
  struct flt *fault;
  int i, j, n;
  if (argc < 2) {
    fprintf(stderr, "usage: %s geom|kmat|kone|dc3d ...\n", argv[0]);
    return 1;
  }
  if (strcmp(argv[1], "geom") == 0 && argc == 3) {
    load(argv[2], 0, &medium, &fault);
    for (i = 0; i < medium->nrflt; i++) {
      struct flt *f = fault + i;
      printf("%.17e %.17e %.17e %.17e %.17e %.17e %.17e", f->x[0], f->x[1], f->x[2],
	     f->cos_alpha, f->sin_alpha, f->l, f->w);
      for (j = 0; j < 3; j++) printf(" %.17e", f->normal[j]);
      for (j = 0; j < 3; j++) printf(" %.17e", f->t_strike[j]);
      for (j = 0; j < 3; j++) printf(" %.17e", f->t_dip[j]);
      printf("\n");
    }
    return 0;
  }
  if (strcmp(argv[1], "kmat") == 0 && argc == 5) {
    FILE *out;
    double *t;
    int *ir;
    int fs = atoi(argv[3]);
    load(argv[2], fs, &medium, &fault);
    n = medium->nrflt;
    t = (double *)malloc(sizeof(double) * (size_t)n * n * 6);
    ir = (int *)malloc(sizeof(int) * (size_t)n * n * 2);
    for (i = 0; i < n; i++)
      for (j = 0; j < n; j++)
	traction_pair(medium, fault, i, j, fs,
		      (double (*)[3])(t + ((size_t)i * n + j) * 6),
		      ir + ((size_t)i * n + j) * 2);
    out = fopen(argv[4], "wb");
    if (!out) return 2;
    fwrite(t, sizeof(double), (size_t)n * n * 6, out);
    fwrite(ir, sizeof(int), (size_t)n * n * 2, out);
    fclose(out);
    return 0;
  }
  if (strcmp(argv[1], "kone") == 0 && argc == 4) {
    double tr[2][3], b, sum2 = 0.0;
    int irr[2];
    int fs = atoi(argv[3]);
    load(argv[2], fs, &medium, &fault);
    n = medium->nrflt;
    for (i = 0; i < n; i++) {
      b = 0.0;
      for (j = 0; j < n; j++) {
	traction_pair(medium, fault, i, j, fs, tr, irr);
	b += tr[0][0];
      }
      sum2 += b * b;
    }
    printf("%.10e\n", sqrt(sum2));
    return 0;
  }
  if (strcmp(argv[1], "dc3d") == 0) {
    double a[14], u[12];
    int iret, ifull;
    while (scanf("%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf",
		 a, a + 1, a + 2, a + 3, a + 4, a + 5, a + 6, a + 7, a + 8, a + 9,
		 a + 10, a + 11, a + 12, a + 13) == 14) {
      ifull = (int)a[13];
      dc3d(a, a + 1, a + 2, a + 3, a + 4, a + 5, a + 6, a + 7, a + 8, a + 9,
	   a + 10, a + 11, a + 12, u, u + 1, u + 2, u + 3, u + 4, u + 5, u + 6, u + 7,
	   u + 8, u + 9, u + 10, u + 11, &iret, &ifull);
      for (i = 0; i < 12; i++) printf("%.17e ", u[i]);
      printf("%d\n", iret);
    }
    return 0;
  }
  fprintf(stderr, "%s: bad arguments\n", argv[0]);
  return 1;
}
