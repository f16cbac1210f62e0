/*
 * ora_input.c -- oracle transcription of the reference input grammar and data readers.
 * TEST INFRASTRUCTURE ONLY (see tild_oracle.h).
 *
 *   parse_deck      read_input.cu:12-533 (keyword loop), set_defaults :536-562
 *   read_data_file  config_utils.cu:5-183 (read_config, FP64 fscanf then cast) and
 *                   config_utils.cu:185-364 (read_charge_config, FP32 fscanf)
 *   read_resume     config_utils.cu:405-454
 *   groups          group.cu:30-60 (make_group_all / make_group_type)
 *   integrators     integrator.cu:44-70
 *
 * Parity status: pinned by the example decks under examples/ (header counts, box,
 * positions vs examples/Coacervate/traj.lammpstrj, bond topology rebuilt from the
 * bond section) and by the reference-data step-0 values (examples/Coacervate/data.dat),
 * which depend on every field read here.
 *
 * Deviations (each is undefined behaviour or a process exit in the reference):
 *   - errors return a message instead of die()/exit(1) (die.cu:3-6);
 *   - more than MAX_BONDS bonds on one site is an error (config_utils.cu:117-123
 *     writes past the array);
 *   - out-of-range type/bond indices are errors (the reference indexes out of bounds).
 */
#include "ora_internal.h"
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ERR(...) do { snprintf(err, errlen, __VA_ARGS__); return -1; } while (0)

/* ---------- minimal istringstream >> emulation over one line ---------- */
typedef struct { const char* p; const char* end; } wstream;
static int next_word(wstream* w, char* out, int cap) {
    while (w->p < w->end && isspace((unsigned char)*w->p)) w->p++;
    if (w->p >= w->end) return 0;
    int n = 0;
    while (w->p < w->end && !isspace((unsigned char)*w->p)) {
        if (n < cap - 1) out[n++] = *w->p;
        w->p++;
    }
    out[n] = 0;
    return 1;
}
static int to_int(const char* s, int* v) {      /* std::stoi: leading integer */
    char* e; long r = strtol(s, &e, 10);
    if (e == s) return 0;
    *v = (int)r; return 1;
}
static int to_float(const char* s, float* v) {  /* std::stof */
    char* e; float r = strtof(s, &e);
    if (e == s) return 0;
    *v = r; return 1;
}

/* line iterator over the deck text (std::getline semantics) */
typedef struct { const char* p; const char* end; int eof; } lines;
static int get_line(lines* L, const char** b, const char** e) {
    if (L->p >= L->end) { L->eof = 1; *b = *e = L->end; return 0; }
    const char* s = L->p;
    const char* nl = memchr(s, '\n', (size_t)(L->end - s));
    if (!nl) { *b = s; *e = L->end; L->p = L->end; L->eof = 1; return 1; }
    *b = s; *e = nl; L->p = nl + 1;
    if (L->p >= L->end) L->eof = 1;
    return 1;
}

static void join_path(char* out, size_t cap, const char* base, const char* name) {
    if (name[0] == '/' || !base || !base[0]) snprintf(out, cap, "%s", name);
    else snprintf(out, cap, "%s/%s", base, name);
}

static int add_group(ora_sys* s, const char* name) {
    ora_group* g = (ora_group*)realloc(s->grp, sizeof(ora_group) * (s->ngrp + 1));
    if (!g) return -1;
    s->grp = g;
    memset(&s->grp[s->ngrp], 0, sizeof(ora_group));
    snprintf(s->grp[s->ngrp].name, sizeof(s->grp[s->ngrp].name), "%s", name);
    s->grp[s->ngrp].index = (int*)calloc((size_t)(s->ns > 0 ? s->ns : 1), sizeof(int));
    return s->ngrp++;
}

/* config_utils.cu:5-183 / :185-364 */
static int read_data_file(ora_sys* s, const char* path, int charged, char* err, int errlen) {
    FILE* inp = fopen(path, "r");
    if (!inp) ERR("Failed to open %s!", path);
    char tt[120];
    int i, di, ind, ltp;
    int ok = 1;
#define FG() ((void)!fgets(tt, 120, inp))
    FG(); FG();
    ok &= fscanf(inp, "%d", &s->ns) == 1; FG();
    ok &= fscanf(inp, "%d", &s->n_total_bonds) == 1; FG();
    ok &= fscanf(inp, "%d", &s->n_total_angles) == 1; FG();
    FG();
    ok &= fscanf(inp, "%d", &s->ntypes) == 1; FG();
    ok &= fscanf(inp, "%d", &s->nbond_types) == 1; FG();
    ok &= fscanf(inp, "%d", &s->nangle_types) == 1; FG();
    if (!ok || s->ns <= 0 || s->ntypes <= 0) { fclose(inp); ERR("malformed data header in %s", path); }
    FG();
    for (int d = 0; d < 3; ++d) {
        if (!charged) {
            double lo, hi;
            ok &= fscanf(inp, "%lf %lf", &lo, &hi) == 2; FG();
            if (d < 2 || s->dim > 2) s->Lf[d] = (float)(hi - lo);
        } else {
            float lo, hi;
            ok &= fscanf(inp, "%f %f", &lo, &hi) == 2; FG();
            if (d < 2 || s->dim > 2) s->Lf[d] = (float)(hi - lo);
        }
    }
    if (s->dim <= 2) s->Lf[2] = 1.0f;
    s->Vf = 1.0f;
    for (int d = 0; d < s->dim; ++d) { s->Lhf[d] = 0.5f * s->Lf[d]; s->Vf *= s->Lf[d]; }

    int ns = s->ns, D = s->dim;
    s->x = (double*)calloc((size_t)ns * D, sizeof(double));
    s->xo = (double*)calloc((size_t)ns * D, sizeof(double));
    s->tp = (int*)calloc(ns, sizeof(int));
    s->molec = (int*)calloc(ns, sizeof(int));
    s->q = (double*)calloc(ns, sizeof(double));
    s->massf = (float*)calloc(s->ntypes, sizeof(float));
    s->difff = (float*)calloc(s->ntypes, sizeof(float));
    for (i = 0; i < s->ntypes; ++i) s->difff[i] = 1.0f;   /* initialize.cu:312-313 */
    s->n_bonds = (int*)calloc(ns, sizeof(int));
    s->bonded_to = (int*)calloc((size_t)ns * ORA_MAX_BONDS, sizeof(int));
    s->bond_type = (int*)calloc((size_t)ns * ORA_MAX_BONDS, sizeof(int));
    int nbt = s->nbond_types > 0 ? s->nbond_types : 1;
    s->bond_kf = (float*)calloc(nbt, sizeof(float));
    s->bond_reqf = (float*)calloc(nbt, sizeof(float));

    FG();
    FG(); FG();
    for (i = 0; i < s->ntypes; ++i) {
        if (!charged) { double m; ok &= fscanf(inp, "%d %lf", &di, &m) == 2; FG();
            if (di < 1 || di > s->ntypes) { fclose(inp); ERR("bad mass type %d", di); }
            s->massf[di - 1] = (float)m; }
        else { float m; ok &= fscanf(inp, "%d %f", &di, &m) == 2; FG();
            if (di < 1 || di > s->ntypes) { fclose(inp); ERR("bad mass type %d", di); }
            s->massf[di - 1] = m; }
    }
    FG();
    FG(); FG();
    for (i = 0; i < ns; ++i) {
        if (feof(inp)) { fclose(inp); ERR("Premature end of input.conf!"); }
        if (fscanf(inp, "%d %d %d", &ind, &di, &ltp) != 3) { fclose(inp); ERR("malformed atom line %d", i + 1); }
        ind -= 1;
        if (ind < 0 || ind >= ns || ltp < 1 || ltp > s->ntypes) { fclose(inp); ERR("bad atom id/type at line %d", i + 1); }
        s->molec[ind] = di - 1;
        s->tp[ind] = ltp - 1;
        if (charged) { float qv; ok &= fscanf(inp, "%f", &qv) == 1; s->q[ind] = qv; }
        for (int j = 0; j < D; ++j) {
            if (!charged) { double v; ok &= fscanf(inp, "%lf", &v) == 1; s->x[ind * D + j] = (double)(float)v; }
            else { float v; ok &= fscanf(inp, "%f", &v) == 1; s->x[ind * D + j] = v; }
        }
        FG();
    }
    FG();
    FG(); FG();
    for (i = 0; i < s->n_total_bonds; ++i) {
        if (feof(inp)) { fclose(inp); ERR("Premature end of input.conf!"); }
        int a, bt, i1, i2;
        if (fscanf(inp, "%d", &a) != 1 || fscanf(inp, "%d", &bt) != 1 ||
            fscanf(inp, "%d", &i1) != 1 || fscanf(inp, "%d", &i2) != 1) { fclose(inp); ERR("malformed bond %d", i + 1); }
        int b_type = bt - 1; i1 -= 1; i2 -= 1;
        if (i2 < i1) { int t = i2; i2 = i1; i1 = t; }
        if (i1 < 0 || i2 >= ns || b_type < 0 || b_type >= nbt) { fclose(inp); ERR("bad bond %d", i + 1); }
        if (s->n_bonds[i1] >= ORA_MAX_BONDS || s->n_bonds[i2] >= ORA_MAX_BONDS) { fclose(inp); ERR("more than %d bonds on a site", ORA_MAX_BONDS); }
        s->bonded_to[i1 * ORA_MAX_BONDS + s->n_bonds[i1]] = i2;
        s->bond_type[i1 * ORA_MAX_BONDS + s->n_bonds[i1]] = b_type;
        s->n_bonds[i1]++;
        s->bonded_to[i2 * ORA_MAX_BONDS + s->n_bonds[i2]] = i1;
        s->bond_type[i2 * ORA_MAX_BONDS + s->n_bonds[i2]] = b_type;
        s->n_bonds[i2]++;
    }
    fclose(inp);
#undef FG
    if (!ok) ERR("malformed data file %s", path);
    if (s->n_total_angles > 0) ERR("Angles not yet set up on device!");   /* initialize.cu:166-168 */
    return 0;
}

/* config_utils.cu:405-454 */
static int read_resume(ora_sys* s, const char* path, char* err, int errlen) {
    FILE* inp = fopen(path, "r");
    if (!inp) ERR("Failed to open %s!", path);
    char tt[120];
    int ns_tmp;
    (void)!fgets(tt, 100, inp); (void)!fgets(tt, 100, inp); (void)!fgets(tt, 100, inp);
    if (fscanf(inp, "%d", &ns_tmp) != 1 || ns_tmp != s->ns) { fclose(inp); ERR("number of sites in resume file doesn't match input!"); }
    for (int k = 0; k < 6; ++k) (void)!fgets(tt, 100, inp);
    for (int i = 0; i < s->ns; ++i) {
        int t1, t2, t3;
        if (fscanf(inp, "%d %d %d  ", &t1, &t2, &t3) != 3) { fclose(inp); ERR("malformed resume line %d", i + 1); }
        for (int j = 0; j < s->dim; ++j) { float v; (void)!fscanf(inp, "%f ", &v); s->x[i * s->dim + j] = v; }
        for (int j = s->dim; j < 3; ++j) { float f1; (void)!fscanf(inp, "%f", &f1); }
        if (s->do_charges) { float f1; (void)!fscanf(inp, "%f", &f1); }
        (void)!fgets(tt, 5, inp);
    }
    fclose(inp);
    return 0;
}

static int style_line(ora_sys* s, lines* L, const char* want, int kind, char* err, int errlen) {
    const char *b, *e;
    get_line(L, &b, &e);
    wstream w = {b, e};
    char tok[256];
    tok[0] = 0;
    next_word(&w, tok, sizeof tok);
    if (strcmp(tok, want) != 0) ERR("Read: %s but expected '%s'", tok, want);
    ora_style st;
    memset(&st, 0, sizeof st);
    st.kind = kind;
    int t1, t2; float A, a, b2;
    if (!next_word(&w, tok, sizeof tok) || !to_int(tok, &t1)) ERR("bad %s line", want);
    if (!next_word(&w, tok, sizeof tok) || !to_int(tok, &t2)) ERR("bad %s line", want);
    if (!next_word(&w, tok, sizeof tok) || !to_float(tok, &A)) ERR("bad %s line", want);
    st.t1 = t1 - 1; st.t2 = t2 - 1; st.A = A; st.A_final = A;
    if (kind == ORA_GAUSS) {                           /* read_input.cu:319-331 */
        if (!next_word(&w, tok, sizeof tok) || !to_float(tok, &a)) ERR("bad gaussian line");
        st.sigma2 = (double)a; st.sigma2 *= st.sigma2;
    } else if (kind == ORA_ERF || kind == ORA_GAUSS_ERF) {  /* read_input.cu:195-210, 376-391 */
        if (!next_word(&w, tok, sizeof tok) || !to_float(tok, &a)) ERR("bad %s line", want);
        if (!next_word(&w, tok, sizeof tok) || !to_float(tok, &b2)) ERR("bad %s line", want);
        st.Rp = (double)a; st.sigma2 = (double)b2; st.sigma2 *= st.sigma2;
    } else {                                           /* read_input.cu:256-273 */
        int ph, dr, np;
        if (!next_word(&w, tok, sizeof tok) || !to_int(tok, &ph)) ERR("bad fieldphase line");
        if (!next_word(&w, tok, sizeof tok) || !to_int(tok, &dr)) ERR("bad fieldphase line");
        if (!next_word(&w, tok, sizeof tok) || !to_int(tok, &np)) ERR("bad fieldphase line");
        st.phase = ph; st.dir = dr; st.n_periods = np;
    }
    if (st.t1 >= s->ntypes || st.t2 >= s->ntypes || st.t1 < 0 || st.t2 < 0)
        ERR("Invalid particle type in %s input!", want);
    if (next_word(&w, tok, sizeof tok) && strcmp(tok, "ramp") == 0) {
        float fa;
        if (!next_word(&w, tok, sizeof tok) || !to_float(tok, &fa)) ERR("bad ramp value");
        st.ramp = 1; st.A_final = fa;
    }
    ora_style* n = (ora_style*)realloc(s->st, sizeof(ora_style) * (s->nst + 1));
    if (!n) ERR("oom");
    s->st = n; s->st[s->nst++] = st;
    return 0;
}

int ora_parse(ora_sys* s, const char* text, const char* base_dir, char* err, int errlen) {
    /* set_defaults, read_input.cu:536-562 */
    s->dim = 2; s->delt_f = 0.005f; s->pmeorder = 1;
    s->Nx[0] = s->Nx[1] = 35; s->Nx[2] = 0;
    s->threads = 512; s->noise_mag_f = sqrtf(2.0f * s->delt_f);
    s->max_steps = 100000; s->log_freq = 2500; s->grid_freq = 0; s->traj_freq = 0;
    s->struc_freq = 0; s->skip_steps = 20000; s->bin_freq = 5000;
    int config_read = 0, read_charge = 0;
    lines L = {text, text + strlen(text), 0};
    const char *b, *e;
    char w1[256], tok[256];
    while (!L.eof) {
        if (!get_line(&L, &b, &e)) break;
        if (e == b || *b == '#') continue;
        wstream w = {b, e};
        while (next_word(&w, w1, sizeof w1)) {
            int iv; float fv;
#define NEED_INT(v) do { if (!next_word(&w, tok, sizeof tok) || !to_int(tok, &(v))) ERR("missing integer after %s", w1); } while (0)
#define NEED_FLT(v) do { if (!next_word(&w, tok, sizeof tok) || !to_float(tok, &(v))) ERR("missing number after %s", w1); } while (0)
            if (!strcmp(w1, "binary_freq")) { NEED_INT(iv); s->bin_freq = iv; }
            else if (!strcmp(w1, "bond")) {
                if (!config_read) ERR("Error in input file, bond keyword before config read");
                int bt; float k, r0;
                NEED_INT(bt);
                if (!next_word(&w, tok, sizeof tok) || strcmp(tok, "harmonic")) ERR("You fool! That bond type is not supported!");
                NEED_FLT(k); NEED_FLT(r0);
                if (bt < 1 || bt > (s->nbond_types > 0 ? s->nbond_types : 1)) ERR("bond type %d out of range", bt);
                s->bond_kf[bt - 1] = k; s->bond_reqf[bt - 1] = r0;
            }
            else if (!strcmp(w1, "charges")) {
                if (config_read) ERR("charges keyword must come before read_data command!");
                read_charge = 1;
                NEED_FLT(fv); s->bjerrum_f = fv;
                NEED_FLT(fv); s->smear_f = fv;
                s->do_charges = 1;
            }
            else if (!strcmp(w1, "compute")) { while (next_word(&w, tok, sizeof tok)) {} }
            else if (!strcmp(w1, "delt")) { NEED_FLT(fv); s->delt_f = fv; s->noise_mag_f = sqrtf(2.f * s->delt_f); }
            else if (!strcmp(w1, "diffusivity")) {
                if (!config_read) ERR("Error in input file, diffusivities defined before config read");
                NEED_INT(iv);
                if (iv > s->ntypes || iv < 1) ERR("Invalid type for diffusivity!");
                NEED_FLT(fv); s->difff[iv - 1] = fv;
            }
            else if (!strcmp(w1, "Dim")) { NEED_INT(iv); s->dim = iv; if (iv != 2 && iv != 3) ERR("Dim must be 2 or 3"); }
            else if (!strcmp(w1, "grid_freq")) { NEED_INT(iv); s->grid_freq = iv; }
            else if (!strcmp(w1, "group")) {
                if (!config_read) ERR("Must read the configuration file before defining a group!");
                char name[64], kind[64];
                if (!next_word(&w, name, sizeof name)) ERR("group needs a name");
                if (!next_word(&w, kind, sizeof kind)) ERR("group needs a style");
                if (strcmp(kind, "type")) ERR("ERROR: Group type %s not supported!", kind);
                NEED_INT(iv);
                int g = add_group(s, name);
                if (g < 0) ERR("oom");
                int t = iv - 1, n = 0;
                for (int i = 0; i < s->ns; ++i) if (s->tp[i] == t) s->grp[g].index[n++] = i;
                s->grp[g].nsites = n;
            }
            else if (!strcmp(w1, "integrator")) {
                if (!config_read) ERR("Integrator must be defined after read_data command!");
                char gname[64], iname[64];
                if (!next_word(&w, gname, sizeof gname) || !next_word(&w, iname, sizeof iname)) ERR("integrator needs group and name");
                int kind;
                if (!strcmp(iname, "EM")) kind = 0; else if (!strcmp(iname, "GJF")) kind = 1;
                else ERR("%s is not a valid integrator!", iname);
                int gi = -1;
                for (int g = 0; g < s->ngrp; ++g) if (!strcmp(s->grp[g].name, gname)) gi = g;
                if (gi < 0) ERR("Group %s not found to integrate!", gname);
                ora_integ* n = (ora_integ*)realloc(s->integ, sizeof(ora_integ) * (s->ninteg + 1));
                if (!n) ERR("oom");
                s->integ = n; s->integ[s->ninteg].kind = kind; s->integ[s->ninteg].group = gi; s->ninteg++;
            }
            else if (!strcmp(w1, "log_freq")) { NEED_INT(iv); s->log_freq = iv; }
            else if (!strcmp(w1, "max_steps")) { NEED_INT(iv); s->max_steps = iv; }
            else if (!strcmp(w1, "n_erfs") || !strcmp(w1, "n_gaussians") ||
                     !strcmp(w1, "n_gaussian_erfs") || !strcmp(w1, "n_fieldphases")) {
                NEED_INT(iv);
                const char* want; int kind;
                if (!strcmp(w1, "n_erfs")) { want = "erf"; kind = ORA_ERF; }
                else if (!strcmp(w1, "n_gaussians")) { want = "gaussian"; kind = ORA_GAUSS; }
                else if (!strcmp(w1, "n_gaussian_erfs")) { want = "gaussian_erf"; kind = ORA_GAUSS_ERF; }
                else { want = "fieldphase"; kind = ORA_FIELDPHASE; }
                /* Erfs.resize(n) replaces any earlier block of the same kind */
                int keep = 0;
                for (int k = 0; k < s->nst; ++k) if (s->st[k].kind != kind) s->st[keep++] = s->st[k];
                s->nst = keep;
                for (int k = 0; k < iv; ++k)
                    if (style_line(s, &L, want, kind, err, errlen)) return -1;
            }
            else if (!strcmp(w1, "Nx")) { NEED_INT(iv); s->Nx[0] = iv; }
            else if (!strcmp(w1, "Ny")) { NEED_INT(iv); s->Nx[1] = iv; }
            else if (!strcmp(w1, "Nz")) { NEED_INT(iv); if (s->dim > 2) s->Nx[2] = iv; }
            else if (!strcmp(w1, "pmeorder")) { NEED_INT(iv); s->pmeorder = iv; if (iv > 4 || iv < 0) ERR("PMEORDER greater than 4 not supported!"); }
            else if (!strcmp(w1, "rand_seed")) { NEED_INT(iv); s->seed = (uint32_t)iv; }
            else if (!strcmp(w1, "read_data")) {
                if (!next_word(&w, tok, sizeof tok)) ERR("read_data needs a file");
                char path[4096];
                join_path(path, sizeof path, base_dir, tok);
                if (read_data_file(s, path, read_charge, err, errlen)) return -1;
                config_read = 1;
                s->ngrp = 0;
                int g = add_group(s, "all");                  /* group.cu:30-42 */
                for (int i = 0; i < s->ns; ++i) s->grp[g].index[i] = i;
                s->grp[g].nsites = s->ns;
            }
            else if (!strcmp(w1, "read_resume")) {
                if (!next_word(&w, tok, sizeof tok)) ERR("read_resume needs a file");
                if (!config_read) ERR("read_resume before read_data");
                char path[4096];
                join_path(path, sizeof path, base_dir, tok);
                if (read_resume(s, path, err, errlen)) return -1;
            }
            else if (!strcmp(w1, "skip_steps")) { NEED_INT(iv); s->skip_steps = iv; }
            else if (!strcmp(w1, "struc_freq")) { NEED_INT(iv); s->struc_freq = iv; }
            else if (!strcmp(w1, "threads")) { NEED_INT(iv); s->threads = iv; }
            else if (!strcmp(w1, "traj_freq")) { NEED_INT(iv); s->traj_freq = iv; }
            else if (!strcmp(w1, "traj_name")) { (void)next_word(&w, tok, sizeof tok); }
            else ERR("Invalid keyword: %s", w1);
#undef NEED_INT
#undef NEED_FLT
        }
    }
    if (!config_read) ERR("no read_data command in input");
    return 0;
}
