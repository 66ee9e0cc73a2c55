/* Plain-data types shared by the wap C ABI (include/wap.h).
 *
 * Memory layout everywhere is planar x-fastest, the layout of the volume images the
 * warp-and-project loop works on: voxel (x, y, z, c) of a W x H x D x C field lives at
 * x + W * (y + H * (z + D * c)). Detector images are u-fastest: pixel (u, v) at
 * u + Nu * v.
 */
#ifndef WAP_TYPES_H
#define WAP_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Scan description used to build per-view matrices. The cone-beam form follows the
 * reference camera: Scene::_calculateCamera / _calculateMatrix
 * (Project/Scene.cpp:65-189), TypicalRayProducer::_calculateMatrix
 * (Project/TypicalRayProducer.cpp:157-232), PSART ctor
 * (Project/WaP_Tomo.cpp:228-256) and PSART::RotateVolume (Project/WaP_Tomo.cpp:1764-1797).
 * The parallel-beam form (parallel = 1) is an extension with the same rotation and
 * detector conventions and rays along the view axis. */
typedef struct WapScanGeometry {
  int vol_dims[3];       /* W, H, D voxels */
  float vol_spacing;     /* voxel pitch (reference: spacings[1]) */
  int det_u, det_v;      /* detector pixels Nu, Nv */
  double det_pitch;      /* detector pixel pitch (reference: m_ds) */
  float sid;             /* source to rotation centre (reference: m_sid) */
  float sdd;             /* source to detector (reference: m_sdd, also the near plane) */
  float translate[3];    /* view translation (reference: m_Translate) */
  int counterclockwise;  /* reference IS_COUNTERCLOCK */
  int parallel;          /* 0 = cone beam (reference), 1 = parallel beam */
  int pad_;
} WapScanGeometry;

/* Everything a projector or backprojector kernel needs for one view. All matrices are
 * column-major 4x4 floats (ele[col*4 + row]) exactly as the reference Matrix4x4. */
typedef struct WapView {
  float p2m[16];    /* pixel -> model (cone: inverse of ViewToPixels*Proj*ModelView;
                       parallel: pixel -> point on the source plane) */
  float m2v[16];    /* model -> voxel */
  float mvv2m[16];  /* ModelView * VolumeToModel, used by the voxel-driven backprojector */
  float eye[4];     /* cone: source position in model space (P2M column 2 / P2M[11]) */
  float dir[4];     /* parallel: unit ray direction in model space */
  double half_w;    /* (Nu-1)*ds/2 (reference m_halfFrustumWidth) */
  double half_h;    /* (Nv-1)*ds/2 */
  double ds;        /* detector pitch */
  float sdd;        /* source-detector distance */
  int parallel;     /* 0 cone, 1 parallel */
} WapView;

/* One warp-and-project outer iteration (PSART::Render, Project/WaP_Tomo.cpp:473-877):
 * SART-prox sweeps of every state, the optical flow of every adjacent pair, and a
 * warp-and-project pass over every state's view window. */
typedef struct WapLoopConfig {
  int dims[3];        /* W, H, D */
  int nstates;        /* deformation states (frames) */
  int nviews;         /* projections */
  int nu, nv;         /* detector */
  float sd;           /* ray step (reference: sqrt(3)/2 * spacing, WaP_Tomo.cpp:380) */
  double alpha;       /* SART relaxation (reference m_alpha, -a) */
  float gmax;         /* volume upper clamp (reference G_Max) */
  int nb_sart;        /* SART sweeps per state (reference nb_proxiSART) */
  int roi[8];         /* m_roi: box margins x0,x1,y0,y1,z0,z1 + corner-cut sizes */
  int cut[8];         /* flow ROI margins at full resolution (reference cut[]) */
  float flow_tol, flow_lambda, flow_eta, flow_sigma;
  int flow_iters, flow_warps, flow_scales;
  int do_flow, do_warp;
  float timeunit;     /* views per state, for the warp time offset */
} WapLoopConfig;

/* Settings of the reconstruction driver's command line (Project/main.cpp:15-143). The
 * reference parses them with argtable2. Flags and defaults are the same. Options marked
 * "required" have a minimum count of 1 in the reference's table. Fields the loop does not
 * read are parsed and kept, as there. */
#define WAP_CLI_PATH 1024
#define WAP_CLI_MAXFILES 20
typedef struct WapCliConfig {
  double sigma;          /* -s/--Sigma: TV weight TV_w (0.3) */
  double tau;            /* -t/--Tau: temporal-prior weight TP_w (0.3) */
  double lambda;         /* -l/--Lambda (0.1; not read by the loop) */
  double alpha;          /* -a/--Alpha: SART relaxation (0.1) */
  double voxel_spacing;  /* -c/--voxelspacing (1.0) */
  double det_spacing;    /* -d (1.0) */
  double sid;            /* -g/--sid: source to rotation centre (400) */
  double sdd;            /* -j/--sdd: source to detector (200) */
  double oxyz[3];        /* --oxyz: volume centre offset, required, 1..3 values (rest 0) */
  double start_deg[WAP_CLI_MAXFILES]; /* --sdg: required, 1..20 values (InitViews only) */
  int n_start_deg;
  int vol[3];            /* -f/--XYZ: volume W H D, required, 1..3 values (rest 63) */
  int img[2];            /* -u/--imgUV: detector W H, required, 1..2 values (rest 512) */
  int nframes;           /* -v/--nframes (1) */
  int nrounds;           /* -r (1) */
  int algo_iters;        /* -b/--AlgorithmIter (2; not read by the loop) */
  int sart_iters;        /* -p/--SartIter (2; not read by the loop) */
  int total_proj;        /* projections: 192, fixed in the reference driver; --nproj */
  int n_proj_files;      /* -i/--projsvolume: required, 1..20 files; the first is read */
  char proj_files[WAP_CLI_MAXFILES][WAP_CLI_PATH];
  char output[WAP_CLI_PATH]; /* -o: output prefix ("Test/test") */
  char prior[16];        /* --prior {STV,SAD,ATV} (ATV; not read by the loop) */
  char bp[16];           /* --bp {Voxelbased,Raybased} (Voxelbased) */
  /* additions, no reference flag */
  char angles[WAP_CLI_PATH]; /* --angles: angle list ("ours_foam_compress.txt") */
  int lds[3];            /* --lds Ns,p,Nz: low-discrepancy angles instead of the list */
  int use_lds;
  int save;              /* --save {none,final,all} (final) */
  int device;            /* --device (0) */
} WapCliConfig;

/* Storage element type of volumes, flows and sinograms. */
typedef enum WapDType { WAP_F32 = 0, WAP_BF16 = 1, WAP_F16 = 2 } WapDType;

#ifdef __cplusplus
}
#endif

#endif /* WAP_TYPES_H */
