/*
 * dsaX_beamformer -- drop-in command line for the MI355X beamformer.
 *
 * Same flags and defaults as the reference (src/dsaX_beamformer.cu:622-638,
 * 684-789): -c core -d -f weights -i key -o key -z fch1 -a flagants -s -q -g
 * -t -h.  The per-block loop follows :939-1090: weights -> beamform -> AGC ->
 * write, first block zeroed, scale applied to the next block, test pattern
 * (blockct%128)+32, ANTPOL_RMS every 20 blocks.
 *
 * Block I/O.  The reference reads psrdada ring -i and writes ring -o.  This
 * build has no psrdada library, so blocks come from a file or stdin (-I) and
 * go to a file or stdout (-O); a leading DADA ASCII header (HDR_SIZE bytes,
 * 4096 by default) is copied to the output verbatim, as the reference copies
 * the ring header (:839-866).  A psrdada binding replaces read_full/fwrite
 * with the reference's ipcio_open_block_read / ipcio_write sequence
 * (:814-866, :942, :1071, :1088) around the same library calls.
 *
 * Extra options (reference defaults when absent):
 *   -I file   input blocks ("-" = stdin)      -O file  output blocks ("-" = stdout)
 *   -n nant   antennas per block (63)          -T tavg  samples per output (8)
 *   -P prec   fp16 | fp8split | fp6split | fp6split3 (fp16)   -W mode  group | fine (group)
 *   -A mode   theta | sin (theta)              -r       input is dsaX_split -r order
 *   -N n      stop after n blocks              -w       re-read weights every block
 *   -B        beam-major output [beam][t][ch]: each 64-beam slice is the buffer
 *             dsaX_dbnic's transmit thread builds by its transpose
 *             (dsaX_dbnic.c:111-116), so the corner turn needs none
 */
#define _GNU_SOURCE
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <syslog.h>
#include <unistd.h>

#include "dsaxbf.h"

static void usage(void) {
  fprintf(stdout,
          "dsaX_beamformer [options]\n"
          " -c core   bind process to CPU core [no default]\n"
          " -d send debug messages to syslog\n"
          " -f filename for antenna stuff [no default]\n"
          " -i input key [default REORDER_BLOCK_KEY2]\n"
          " -o output key [default BF_BLOCK_KEY]\n"
          " -z fch1 in MHz [default 1530]\n"
          " -a flagants file\n"
          " -s stuffants \n"
          " -q do incoherent beam \n"
          " -g skip AGC \n"
          " -t test pattern \n"
          " -h print usage\n"
          " -I infile (- = stdin)   -O outfile (- = stdout)   -n nant [63]\n"
          " -T tavg [8]  -P fp16|fp8split|fp6split|fp6split3  -W group|fine  -A theta|sin  -r  -N nblocks\n"
          " -w  -B (beam-major output)\n");
}

/* Read exactly n bytes unless EOF; returns bytes read. */
static size_t read_full(FILE *f, void *buf, size_t n) {
  size_t got = 0;
  while (got < n) {
    size_t r = fread((char *)buf + got, 1, n - got, f);
    if (r == 0) break;
    got += r;
  }
  return got;
}

int main(int argc, char *argv[]) {
  openlog("dsaX_beamformer", LOG_CONS | LOG_PID | LOG_NDELAY, LOG_LOCAL0);
  syslog(LOG_NOTICE, "Program started by User %d", getuid());

  int core = -1, debug = 0, stuffants = 0, test_pattern = 0, agc_on = 1;
  int nant = 63, tavg = 8, prec = DSAXBF_PREC_FP16, wmode = DSAXBF_WEIGHTS_GROUP;
  int amode = DSAXBF_ANGLE_THETA, order = DSAXBF_ORDER_CAPTURE, reload = 0;
  int layout = DSAXBF_LAYOUT_TIME_MAJOR;
  long max_blocks = -1;
  unsigned in_key = DSAXBF_REORDER_BLOCK_KEY2, out_key = DSAXBF_BF_BLOCK_KEY;
  float fch1 = 1530.0f;
  const char *fnam = "nofile", *flagants = "nofile", *infile = NULL, *outfile = NULL;
  int arg;
  while ((arg = getopt(argc, argv, "c:f:i:o:z:a:tsqdghI:O:n:T:P:W:A:rN:wB")) != -1) {
    switch (arg) {
      case 'c': core = atoi(optarg); break;
      case 'i':
        if (sscanf(optarg, "%x", &in_key) != 1) {
          syslog(LOG_ERR, "could not parse key from %s", optarg);
          return EXIT_FAILURE;
        }
        break;
      case 'o':
        if (sscanf(optarg, "%x", &out_key) != 1) {
          syslog(LOG_ERR, "could not parse key from %s", optarg);
          return EXIT_FAILURE;
        }
        break;
      case 'f': fnam = optarg; break;
      case 'a': flagants = optarg; break;
      case 'z': fch1 = (float)atof(optarg); break;
      case 'd': debug = 1; break;
      case 'g': agc_on = 0; break;
      case 't': test_pattern = 1; break;
      case 's': stuffants = DSAXBF_STUFF_ANTS; break;
      case 'q': stuffants = DSAXBF_STUFF_IB; break;
      case 'h': usage(); return EXIT_SUCCESS;
      case 'I': infile = optarg; break;
      case 'O': outfile = optarg; break;
      case 'n': nant = atoi(optarg); break;
      case 'T': tavg = atoi(optarg); break;
      case 'P':
        if (strcmp(optarg, "fp16") == 0) prec = DSAXBF_PREC_FP16;
        else if (strcmp(optarg, "fp8split") == 0) prec = DSAXBF_PREC_FP8SPLIT;
        else if (strcmp(optarg, "fp6split") == 0) prec = DSAXBF_PREC_FP6SPLIT;
        else if (strcmp(optarg, "fp6split3") == 0) prec = DSAXBF_PREC_FP6SPLIT3;
        else {
          fprintf(stderr, "dsaX_beamformer: unknown precision %s\n", optarg);
          usage();
          return EXIT_FAILURE;
        }
        break;
      case 'W': wmode = strcmp(optarg, "fine") == 0 ? DSAXBF_WEIGHTS_FINE : DSAXBF_WEIGHTS_GROUP; break;
      case 'A': amode = strcmp(optarg, "sin") == 0 ? DSAXBF_ANGLE_SIN : DSAXBF_ANGLE_THETA; break;
      case 'r': order = DSAXBF_ORDER_SPLIT_R; break;
      case 'N': max_blocks = atol(optarg); break;
      case 'w': reload = 1; break;
      case 'B': layout = DSAXBF_LAYOUT_BEAM_MAJOR; break;
      default: usage(); return EXIT_FAILURE;
    }
  }
  if (!infile || !outfile) {
    syslog(LOG_ERR, "psrdada rings %x/%x are not available in this build; use -I and -O", in_key,
           out_key);
    fprintf(stderr, "dsaX_beamformer: no psrdada in this build; give -I infile -O outfile\n");
    return EXIT_FAILURE;
  }
  if (nant < 1 || nant > 64) {
    /* the weights file (init_weights, :559-602) holds positions and
       calibrations of 64 antennas */
    syslog(LOG_ERR, "-n %d: the weights file holds 64 antennas", nant);
    fprintf(stderr, "dsaX_beamformer: -n %d outside 1..64 (the weights file holds 64 antennas)\n",
            nant);
    usage();
    return EXIT_FAILURE;
  }
  if (wmode == DSAXBF_WEIGHTS_FINE) {
    syslog(LOG_ERR, "the weights file holds one weight per 8-channel group; -W fine needs the API");
    return EXIT_FAILURE;
  }
  syslog(LOG_INFO, "Forming 256 beams with sep %g arcmin, fch1 %g", 1.0, fch1);
  syslog(LOG_INFO, "Using calibrations file %s", fnam);
  syslog(LOG_INFO, "Using flagants file %s", flagants);

  if (core >= 0) {
    cpu_set_t set;
    CPU_ZERO(&set);
    CPU_SET(core, &set);
    if (sched_setaffinity(0, sizeof(set), &set) < 0) syslog(LOG_ERR, "failed to bind to core %d", core);
    else syslog(LOG_NOTICE, "bound to core %d", core);
  }

  dsaxbf_config cfg;
  dsaxbf_config_default(&cfg, nant);
  cfg.tavg = tavg;
  cfg.precision = prec;
  cfg.weight_mode = wmode;
  cfg.angle_mode = amode;
  cfg.stuffants = stuffants;
  cfg.input_order = order;
  cfg.output_layout = layout;
  const size_t block_size = dsaxbf_input_bytes(&cfg), block_out = dsaxbf_output_bytes(&cfg);
  syslog(LOG_INFO, "main: have input and output block sizes %zu %zu", block_size, block_out);

  float antpos64[64], *cal64 = malloc(sizeof(float) * 64 * 48 * 4);
  float *cal = malloc(sizeof(float) * (size_t)nant * 48 * 4);
  double freqs[384];
  dsaxbf_reference_freqs(fch1, 384, freqs);

  dsaxbf_plan *plan = NULL;
  int rc = dsaxbf_create(&plan, &cfg);
  if (rc) {
    syslog(LOG_ERR, "could not create beamformer: %s", dsaxbf_last_error());
    fprintf(stderr, "dsaX_beamformer: %s\n", dsaxbf_last_error());
    return EXIT_FAILURE;
  }
  /* weights: file -> normalise -> flag -> device phase kernel */
  int have_file = strcmp(fnam, "nofile") != 0;
#define LOAD_WEIGHTS()                                                                  \
  do {                                                                                  \
    if (have_file) {                                                                    \
      rc = dsaxbf_read_weights_file(fnam, flagants, antpos64, cal64);                   \
      if (rc) {                                                                         \
        syslog(LOG_ERR, "%s", dsaxbf_last_error());                                     \
        fprintf(stderr, "dsaX_beamformer: %s\n", dsaxbf_last_error());                  \
        return EXIT_FAILURE;                                                            \
      }                                                                                 \
    } else {                                                                            \
      memset(antpos64, 0, sizeof(antpos64));                                            \
      memset(cal64, 0, sizeof(float) * 64 * 48 * 4);                                    \
    }                                                                                   \
    memcpy(cal, cal64, sizeof(float) * (size_t)nant * 48 * 4);                          \
    rc = dsaxbf_set_weights(plan, antpos64, cal, freqs, NULL, 1.0);                     \
    if (rc) {                                                                           \
      syslog(LOG_ERR, "weights: %s", dsaxbf_last_error());                              \
      return EXIT_FAILURE;                                                              \
    }                                                                                   \
  } while (0)
  LOAD_WEIGHTS();

  FILE *fin = strcmp(infile, "-") == 0 ? stdin : fopen(infile, "rb");
  FILE *fout = strcmp(outfile, "-") == 0 ? stdout : fopen(outfile, "wb");
  if (!fin || !fout) {
    syslog(LOG_ERR, "cannot open %s or %s", infile, outfile);
    return EXIT_FAILURE;
  }
  /* optional DADA header passthrough */
  {
    char hdr[4096];
    int c0 = fgetc(fin);
    if (c0 != EOF) ungetc(c0, fin);
    if (c0 == 'H') {
      size_t got = read_full(fin, hdr, sizeof(hdr));
      size_t hsize = sizeof(hdr);
      char *p = strstr(hdr, "HDR_SIZE");
      if (p) hsize = strtoul(p + 8, NULL, 10);
      if (got < sizeof(hdr) || hsize < sizeof(hdr)) {
        syslog(LOG_ERR, "could not read next header");
        return EXIT_FAILURE;
      }
      fwrite(hdr, 1, sizeof(hdr), fout);
      for (size_t extra = hsize - sizeof(hdr); extra > 0;) {
        size_t n = extra < sizeof(hdr) ? extra : sizeof(hdr);
        if (read_full(fin, hdr, n) != n) return EXIT_FAILURE;
        fwrite(hdr, 1, n, fout);
        extra -= n;
      }
    }
  }

  /* the reference prints ANTPOL_RMS to stdout because its data goes to a
     ring; here stdout may carry the data blocks, so the lines go to stderr */
  FILE *frms = fout == stdout ? stderr : stdout;
  uint8_t *block = malloc(block_size), *out = malloc(block_out);
  float scale[256], beam_power[256], rms[128];
  memset(scale, 0, sizeof(scale));
  dsaxbf_agc agc;
  if (dsaxbf_agc_init(&agc, 256, 48, cfg.npkt / 16, agc_on) || !block || !out) return EXIT_FAILURE;
  syslog(LOG_INFO, "starting observation");
  long blockct = 0;
  int slow_down = 0;
  for (;;) {
    size_t bytes_read = read_full(fin, block, block_size);
    if (bytes_read == 0) break;
    if (bytes_read < block_size) memset(block + bytes_read, 0, block_size - bytes_read);
    blockct++;
    if (reload) LOAD_WEIGHTS();
    if (slow_down == 0) /* rms_printer: packet 8 of the first four ints, group 24 */
      for (int st = 0; st < 4; st++)
        if (dsaxbf_antpol_rms(&cfg, block, st * 16 + 8, 24, rms) == 0)
          for (int i = 0; i < 128; i++) fprintf(frms, "ANTPOL_RMS %d %d %f\n", i % 64, i / 64, rms[i]);
    rc = dsaxbf_process_block(plan, block, scale, out, NULL, beam_power);
    if (rc) {
      syslog(LOG_ERR, "beamform: %s", dsaxbf_last_error());
      fprintf(stderr, "dsaX_beamformer: %s\n", dsaxbf_last_error());
      return EXIT_FAILURE;
    }
    int zero = 0;
    dsaxbf_agc_update(&agc, beam_power, scale, &zero);
    if (agc.bpctr <= 15 && debug)
      for (int i = 0; i < 256; i++) syslog(LOG_INFO, "coeff %d %d %g", agc.bpctr - 1, i, scale[i]);
    if (zero) memset(out, 0, block_out);
    else if (test_pattern) memset(out, (int)((blockct % 128) + 32), block_out);
    if (fwrite(out, 1, block_out, fout) < block_out) {
      syslog(LOG_ERR, "main: failed to write all data to datablock [output]");
      return EXIT_FAILURE;
    }
    if (debug) syslog(LOG_DEBUG, "written block %ld", blockct - 1);
    if (++slow_down >= 20) slow_down = 0;
    if (bytes_read < block_size) break;   /* EOD */
    if (max_blocks > 0 && blockct >= max_blocks) break;
  }
  fflush(fout);
  if (fout != stdout) fclose(fout);
  if (fin != stdin) fclose(fin);
  dsaxbf_agc_free(&agc);
  dsaxbf_destroy(plan);
  free(block);
  free(out);
  free(cal);
  free(cal64);
  return EXIT_SUCCESS;
}
