// batch.hip -- many independent HMC chains advanced together, one launch per phase over all the
// chains of a group.
//
// The reference runs one chain per process (organic_hmc.c:296-353).  Here the chains of uneven
// lattice and Nt are split into groups of chains that share the arithmetic, the force sweep's
// columns per workgroup (which fix its partial sums), N rounded up to a wave (the site-form
// kernels run one site per thread, on that many threads) and the number of MD steps.
// Every kernel of a trajectory runs once per group with workgroups (x, y, z) = block (x, y) of
// chain z: the argument records of all chains sit in device tables built once (every pointer is
// fixed per context), and a chain's blocks past its own grid return.  Phases whose launch count
// depends on the size (Gauss-Jordan block steps: N/32, force-sweep segments: Nt/32) run as many
// launches as the group's largest chain needs, each over the chains that still have a step.
// The groups are dealt to two streams, the largest estimated work first onto the less loaded
// one; each stream's launch sequence (its groups one after the other) is captured once as a
// hipGraph, and the two overlap.  On 64 chains of L 8..24 two streams ran 1.00e3 trajectories/s,
// against 0.80e3 on one, 0.87e3 on four and 0.93e3 on five (one per group).
//
// The kernels are the single-chain kernels' bodies, with each chain's own grid for every sum,
// so a chain gives the same bits in a batch as alone (ephmc_trajectory), whatever shapes the
// single-chain path picks for its chain and transforms.
#include "internal.h"
#include <cstdlib>
#include <cstring>

using namespace ephmc;

struct ephmc_batch {
	struct Group {
		std::vector<unsigned> idx;        // chains (positions in the batch)
		unsigned threads = 0;             // LDS layout of the site-form kernels (256 / 512 / 1024)
		unsigned block = 0;               // threads launched: N rounded up to a wave
		std::vector<BatchOp> ops;
		std::vector<unsigned> seq;        // the trajectory: op indices in launch order
		double work = 0;                  // sum of N^2 Nt (steps + 1) over its chains
	};
	struct Lane {                         // one stream: its groups' sequences one after the other
		std::vector<unsigned> groups;
		double work = 0;
		hipStream_t stream = nullptr;
		hipEvent_t done = nullptr;
		hipGraph_t graph = nullptr;
		hipGraphExec_t exec = nullptr;
	};
	std::vector<ephmc_ctx *> ch;
	std::vector<hipEvent_t> ch_ev;        // ordering after each context's own earlier work
	std::vector<Group> groups;
	std::vector<Lane> lanes;
	unsigned char *d_tab = nullptr;       // every op's table
	std::vector<size_t> in_off;           // per chain: offset (doubles) of its normals; the uniform follows them
	size_t in_len = 0;
	double *h_in[2] = {nullptr, nullptr}, *d_in = nullptr;
	hipEvent_t in_ev[2] = {nullptr, nullptr};
	int in_next = 0;
	double *d_rec = nullptr, *h_rec = nullptr;   // SCAL_COUNT scalars per chain
	hipStream_t lead = nullptr;           // inputs in, records out
	hipEvent_t up = nullptr, fin = nullptr;
	bool pending = false;
	bool counted = false;                 // batch_refs of the contexts raised
	bool eager = false;                   // EPHMC_BATCH_EAGER=1: plain launches instead of the graphs (profiling)
	int device = 0;
};

namespace {

using Batch = ephmc_batch;

int check_chain(ephmc_ctx *c, unsigned i)
{
	const ephmc_config &g = c->cfg;
	if(!c->chain_ready){ set_error("batch: chain %u has no state (ephmc_chain_init)", i); return EPHMC_ERR_ARG; }
	if(g.no_fourier_acc || g.classical_fourier_acc || g.stabilize || g.single_particle || g.zero_filling || g.constant_force
			|| g.schur_form_inv){
		set_error("batch: chain %u uses a mode the batch does not run (it runs the Fourier-accelerated "
			"trajectory without stabilisation, SINGLE_PARTICLE, ZERO_FILLING, CONSTANT_FORCE or SCHUR_FORM_INV)", i);
		return EPHMC_ERR_UNSUPPORTED;
	}
	if(c->ns <= 32 || c->ns > 1024){ set_error("batch: chain %u has N = %u (the batch runs 33 <= N <= 1024)", i, c->ns); return EPHMC_ERR_UNSUPPORTED; }
	if(g.steps == 0){ set_error("batch: chain %u has no MD steps", i); return EPHMC_ERR_UNSUPPORTED; }
	if(c->poly.deg != 4){ set_error("batch: chain %u has Taylor order %d (the batch runs order 4)", i, c->poly.deg); return EPHMC_ERR_UNSUPPORTED; }
	if(g.precision != EPHMC_FP64 && g.precision != EPHMC_MIXED){ set_error("batch: chain %u: precision %d", i, g.precision); return EPHMC_ERR_ARG; }
	if(c->timing){ set_error("batch: chain %u has timing enabled", i); return EPHMC_ERR_ARG; }
	return EPHMC_OK;
}

unsigned threads_for(unsigned ns){ return ns <= 256 ? 256 : ns <= 512 ? 512 : 1024; }

unsigned push(std::vector<BatchOp> &ops){ ops.emplace_back(); return (unsigned)ops.size() - 1; }

// The launch sequence of one group: trajectory_body (ctx.hip) in the Fourier-accelerated form,
// then the Metropolis decision, the accept copies and the record gather.
int plan_group(Batch &b, Batch::Group &g)
{
	std::vector<BatchOp> &ops = g.ops;
	const unsigned copy_x = push(ops), filter = push(ops), evolve_half = push(ops);
	unsigned ham_old[5], ham_new[5];
	for(int q = 0; q < 5; q++) ham_old[q] = push(ops);      // five consecutive ops (batch_hamilton)
	const unsigned coef_md = push(ops), chain_md = push(ops), trans = push(ops);
	const unsigned gj_first = push(ops), gj_fb = push(ops);
	std::vector<BatchOp> gj_steps, adj_segs;
	const unsigned adj_red = push(ops), evolve_kick = push(ops), evolve_last = push(ops);
	const unsigned coef_ex = push(ops), chain_ex = push(ops), addid = push(ops);
	for(int q = 0; q < 5; q++) ham_new[q] = push(ops);
	const unsigned metro = push(ops), acc_x = push(ops), acc_m = push(ops), rec = push(ops);
	for(unsigned i : g.idx){
		ephmc_ctx *c = b.ch[i];
		const int prec = c->cfg.precision;
		const unsigned steps = c->cfg.steps;
		const double h = c->traj_length / steps;
		void *tfin = prec == EPHMC_FP64 ? (void *)c->d_Tfin : (void *)c->d_Tfin32;
		EPHMC_TRY(batch_copy(ops[copy_x], c->d_x, c->d_xnew, c->dim, nullptr));
		EPHMC_TRY(batch_fa_filter(ops[filter], c, b.d_in + b.in_off[i], c->d_p));
		EPHMC_TRY(batch_hamilton(&ops[ham_old[0]], c, c->d_xnew, c->d_p, SCAL_LOGDET_CHAIN, SCAL_H_OLD));
		EPHMC_TRY(batch_fa_evolve(ops[evolve_half], c, c->d_xnew, c->d_p, 0.5 * h, nullptr, 0));
		EPHMC_TRY(batch_coef(ops[coef_md], c, c->d_xnew, prec != EPHMC_FP64));
		EPHMC_TRY(batch_chain(ops[chain_md], c, prec, g.threads, g.block, tfin, c->d_ck, c->ckpt, 0));
		EPHMC_TRY(batch_transpose(ops[trans], c, prec, tfin, c->d_mhat));
		EPHMC_TRY(batch_gj(ops[gj_first], gj_steps, ops[gj_fb], c, c->d_mhat));
		EPHMC_TRY(batch_adjoint(adj_segs, ops[adj_red], c, prec, g.threads, g.block, c->d_mhat, c->d_pdot));
		EPHMC_TRY(batch_fa_evolve(ops[evolve_kick], c, c->d_xnew, c->d_p, h, c->d_pdot, h));
		EPHMC_TRY(batch_fa_evolve(ops[evolve_last], c, c->d_xnew, c->d_p, 0.5 * h, c->d_pdot, h));
		EPHMC_TRY(batch_coef(ops[coef_ex], c, c->d_xnew, false));
		EPHMC_TRY(batch_chain(ops[chain_ex], c, EPHMC_FP64, g.threads, g.block, c->d_mhat, nullptr, 0, 1));
		EPHMC_TRY(batch_add_identity(ops[addid], c, c->d_mhat));
		EPHMC_TRY(batch_hamilton(&ops[ham_new[0]], c, c->d_xnew, c->d_p, SCAL_LOGDET, SCAL_H_NEW));
		EPHMC_TRY(batch_metropolis(ops[metro], c, b.d_in + b.in_off[i] + c->dim));
		EPHMC_TRY(batch_copy(ops[acc_x], c->d_xnew, c->d_x, c->dim, c->d_scal));
		EPHMC_TRY(batch_copy(ops[acc_m], c->d_mhat, c->d_minv, c->mat, c->d_scal));
		EPHMC_TRY(batch_copy(ops[rec], c->d_scal, b.d_rec + (size_t)i * SCAL_COUNT, SCAL_COUNT, nullptr));
	}
	// the variable-length phases become ops of their own
	std::vector<unsigned> gj_idx, adj_idx;
	for(BatchOp &o : gj_steps){ gj_idx.push_back((unsigned)ops.size()); ops.push_back(std::move(o)); }
	for(BatchOp &o : adj_segs){ adj_idx.push_back((unsigned)ops.size()); ops.push_back(std::move(o)); }
	std::vector<unsigned> &s = g.seq;
	auto gj = [&](){
		s.push_back(gj_first);
		for(unsigned k : gj_idx) s.push_back(k);
		s.push_back(gj_fb);
	};
	s = {copy_x, filter};
	for(unsigned k : ham_old) s.push_back(k);
	s.push_back(evolve_half);
	const unsigned steps = b.ch[g.idx[0]]->cfg.steps;
	for(unsigned st = 1; st <= steps; st++){
		s.push_back(coef_md);
		s.push_back(chain_md);
		s.push_back(trans);
		gj();
		for(unsigned k : adj_idx) s.push_back(k);
		s.push_back(adj_red);
		s.push_back(st == steps ? evolve_last : evolve_kick);
	}
	s.push_back(coef_ex);
	s.push_back(chain_ex);
	s.push_back(addid);
	gj();
	for(unsigned k : ham_new) s.push_back(k);
	s.push_back(metro);
	s.push_back(acc_x);
	s.push_back(acc_m);
	s.push_back(rec);
	return EPHMC_OK;
}

int launch_seq(Batch &b, Batch::Lane &l)
{
	for(unsigned gi : l.groups){
		const Batch::Group &g = b.groups[gi];
		for(unsigned k : g.seq){
			const BatchOp &o = g.ops[k];
			if(o.n == 0) continue;
			void *tab = b.d_tab + o.off;
			void *args[] = {&tab};
			EPHMC_HIP(hipLaunchKernel(o.fn, dim3(o.gx, o.gy, o.n), dim3(o.block), args, o.lds, l.stream));
		}
	}
	return EPHMC_OK;
}

int capture(Batch &b, Batch::Lane &g)
{
	EPHMC_HIP(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
	const int rc = launch_seq(b, g);
	hipGraph_t graph = nullptr;
	const hipError_t e = hipStreamEndCapture(g.stream, &graph);
	if(rc != EPHMC_OK){
		if(graph) (void)hipGraphDestroy(graph);
		return rc;
	}
	EPHMC_HIP(e);
	g.graph = graph;
	EPHMC_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
	return EPHMC_OK;
}

void release(Batch *b)
{
	if(!b) return;
	if(b->pending) (void)hipStreamSynchronize(b->lead);
	for(Batch::Lane &g : b->lanes){
		if(g.stream) (void)hipStreamSynchronize(g.stream);
		if(g.exec) (void)hipGraphExecDestroy(g.exec);
		if(g.graph) (void)hipGraphDestroy(g.graph);
		if(g.done) (void)hipEventDestroy(g.done);
		if(g.stream) (void)hipStreamDestroy(g.stream);
	}
	for(hipEvent_t e : b->ch_ev) if(e) (void)hipEventDestroy(e);
	for(int s = 0; s < 2; s++){
		if(b->in_ev[s]) (void)hipEventDestroy(b->in_ev[s]);
		if(b->h_in[s]) (void)hipHostFree(b->h_in[s]);
	}
	if(b->up) (void)hipEventDestroy(b->up);
	if(b->fin) (void)hipEventDestroy(b->fin);
	if(b->lead) (void)hipStreamDestroy(b->lead);
	if(b->h_rec) (void)hipHostFree(b->h_rec);
	if(b->d_rec) (void)hipFree(b->d_rec);
	if(b->d_in) (void)hipFree(b->d_in);
	if(b->d_tab) (void)hipFree(b->d_tab);
	if(b->counted)
		for(ephmc_ctx *c : b->ch) c->batch_refs--;
	delete b;
}

int create(ephmc_ctx *const *ctxs, unsigned n, Batch *b)
{
	b->ch.assign(ctxs, ctxs + n);
	const char *eager = getenv("EPHMC_BATCH_EAGER");
	b->eager = eager && atoi(eager) == 1;
	for(unsigned i = 0; i < n; i++){
		if(!ctxs[i]){ set_error("batch: chain %u is null", i); return EPHMC_ERR_ARG; }
		if(i == 0) b->device = ctxs[0]->device;
		if(ctxs[i]->device != b->device){ set_error("batch: chains on two devices"); return EPHMC_ERR_ARG; }
		for(unsigned j = 0; j < i; j++)
			if(ctxs[j] == ctxs[i]){ set_error("batch: chain %u appears twice", i); return EPHMC_ERR_ARG; }
		EPHMC_TRY(check_chain(ctxs[i], i));
	}
	EPHMC_HIP(hipSetDevice(b->device));
	for(ephmc_ctx *c : b->ch) c->batch_refs++;
	b->counted = true;
	// four columns per workgroup of the force sweep wherever its LDS allows: with many chains
	// filling the GPU the widest sweep is the fastest (64 chains: 1.06e3 trajectories/s against
	// 0.99e3 with widths tuned chain by chain for a chain alone).  The contexts keep the width, so
	// their own trajectories stay bit for bit those of the batch.
	for(ephmc_ctx *c : b->ch)
		if(adjoint_lds(c->ns, c->cfg.precision, 4) <= 160 * 1024) EPHMC_TRY(set_force_cols(c, 4));
	// inputs: normals then the uniform, per chain
	b->in_off.resize(n);
	for(unsigned i = 0; i < n; i++){
		b->in_off[i] = b->in_len;
		b->in_len += ctxs[i]->dim + 1;
	}
	for(int s = 0; s < 2; s++){
		EPHMC_HIP(hipHostMalloc((void **)&b->h_in[s], b->in_len * sizeof(double), hipHostMallocDefault));
		EPHMC_HIP(hipEventCreateWithFlags(&b->in_ev[s], hipEventDisableTiming));
		EPHMC_HIP(hipEventRecord(b->in_ev[s], nullptr));
	}
	EPHMC_HIP(hipMalloc((void **)&b->d_in, b->in_len * sizeof(double)));
	EPHMC_HIP(hipMalloc((void **)&b->d_rec, (size_t)n * SCAL_COUNT * sizeof(double)));
	EPHMC_HIP(hipHostMalloc((void **)&b->h_rec, (size_t)n * SCAL_COUNT * sizeof(double), hipHostMallocDefault));
	EPHMC_HIP(hipStreamCreateWithFlags(&b->lead, hipStreamNonBlocking));
	EPHMC_HIP(hipEventCreateWithFlags(&b->up, hipEventDisableTiming));
	EPHMC_HIP(hipEventCreateWithFlags(&b->fin, hipEventDisableTiming));
	b->ch_ev.resize(n, nullptr);
	for(unsigned i = 0; i < n; i++) EPHMC_HIP(hipEventCreateWithFlags(&b->ch_ev[i], hipEventDisableTiming));
	// groups (EPHMC_BATCH_BY_NT=1: also one group per Nt, for tuning sweeps)
	const char *by_nt_env = getenv("EPHMC_BATCH_BY_NT");
	const bool by_nt = by_nt_env && atoi(by_nt_env) == 1;
	for(unsigned i = 0; i < n; i++){
		ephmc_ctx *c = ctxs[i];
		const int prec = c->cfg.precision;
		const unsigned thr = threads_for(c->ns), blk = (c->ns + 63) / 64 * 64, cols = adjoint_cols(c, prec);
		Batch::Group *g = nullptr;
		for(Batch::Group &q : b->groups){
			ephmc_ctx *r = ctxs[q.idx[0]];
			if(r->cfg.precision == prec && q.threads == thr && q.block == blk && adjoint_cols(r, prec) == cols
					&& r->cfg.steps == c->cfg.steps && (!by_nt || r->nt == c->nt)){
				g = &q;
				break;
			}
		}
		if(!g){
			b->groups.emplace_back();
			g = &b->groups.back();
			g->threads = thr;
			g->block = blk;
		}
		g->idx.push_back(i);
		g->work += (double)c->ns * c->ns * c->nt * (c->cfg.steps + 1);
	}
	// groups to streams: largest work first, each onto the least loaded stream
	std::vector<unsigned> order(b->groups.size());
	for(unsigned q = 0; q < order.size(); q++) order[q] = q;
	std::stable_sort(order.begin(), order.end(), [&](unsigned x, unsigned y){ return b->groups[x].work > b->groups[y].work; });
	unsigned nlanes = 2;                 // EPHMC_BATCH_STREAMS overrides it for tuning sweeps
	if(const char *e = getenv("EPHMC_BATCH_STREAMS")){
		const int v = atoi(e);
		if(v >= 1 && v <= 16) nlanes = (unsigned)v;
	}
	b->lanes.resize(std::min<size_t>(b->groups.size(), nlanes));
	for(unsigned q : order){
		Batch::Lane *l = &b->lanes[0];
		for(Batch::Lane &m : b->lanes) if(m.work < l->work) l = &m;
		l->groups.push_back(q);
		l->work += b->groups[q].work;
	}
	size_t total = 0;
	for(Batch::Group &g : b->groups){
		EPHMC_TRY(plan_group(*b, g));
		for(BatchOp &o : g.ops){
			o.off = total;
			total += (o.tab.size() + 255) & ~(size_t)255;
		}
	}
	EPHMC_HIP(hipMalloc((void **)&b->d_tab, total ? total : 256));
	std::vector<unsigned char> host(total);
	for(Batch::Group &g : b->groups)
		for(BatchOp &o : g.ops) std::memcpy(host.data() + o.off, o.tab.data(), o.tab.size());
	EPHMC_HIP(hipMemcpy(b->d_tab, host.data(), total, hipMemcpyHostToDevice));
	for(Batch::Lane &g : b->lanes){
		EPHMC_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
		EPHMC_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
		if(!b->eager) EPHMC_TRY(capture(*b, g));
	}
	return EPHMC_OK;
}

// Everything a batch trajectory enqueues after its inputs are staged in h (slot s).
int submit(Batch &b, const double *h, int s)
{
	const unsigned n = (unsigned)b.ch.size();
	EPHMC_HIP(hipMemcpyAsync(b.d_in, h, b.in_len * sizeof(double), hipMemcpyHostToDevice, b.lead));
	EPHMC_HIP(hipEventRecord(b.in_ev[s], b.lead));
	EPHMC_HIP(hipEventRecord(b.up, b.lead));
	for(auto &g : b.lanes){
		EPHMC_HIP(hipStreamWaitEvent(g.stream, b.up, 0));
		if(g.exec) EPHMC_HIP(hipGraphLaunch(g.exec, g.stream));
		else EPHMC_TRY(launch_seq(b, g));
		EPHMC_HIP(hipEventRecord(g.done, g.stream));
		EPHMC_HIP(hipStreamWaitEvent(b.lead, g.done, 0));
	}
	EPHMC_HIP(hipMemcpyAsync(b.h_rec, b.d_rec, (size_t)n * SCAL_COUNT * sizeof(double), hipMemcpyDeviceToHost, b.lead));
	EPHMC_HIP(hipEventRecord(b.fin, b.lead));
	// each context's later work runs after this trajectory
	for(unsigned i = 0; i < n; i++) EPHMC_HIP(hipStreamWaitEvent(b.ch[i]->stream, b.fin, 0));
	return EPHMC_OK;
}

} // namespace

extern "C" {

int ephmc_batch_create(ephmc_ctx *const *ctxs, unsigned n, ephmc_batch **out)
{
	if(!ctxs || !out || n == 0){ set_error("null argument"); return EPHMC_ERR_ARG; }
	*out = nullptr;
	Batch *b = new Batch();
	const int rc = create(ctxs, n, b);
	if(rc != EPHMC_OK){
		release(b);
		return rc;
	}
	*out = b;
	return EPHMC_OK;
}

int ephmc_batch_trajectory(ephmc_batch *b, const double *const *normals, const double *uniforms)
{
	if(!b || !normals || !uniforms){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(b->pending){ set_error("batch: the previous trajectory has not been collected (ephmc_batch_sync)"); return EPHMC_ERR_ARG; }
	const unsigned n = (unsigned)b->ch.size();
	for(unsigned i = 0; i < n; i++)
		if(!normals[i]){ set_error("batch: normals of chain %u are null", i); return EPHMC_ERR_ARG; }
	EPHMC_HIP(hipSetDevice(b->device));
	const int s = b->in_next;
	b->in_next ^= 1;
	EPHMC_HIP(hipEventSynchronize(b->in_ev[s]));
	double *h = b->h_in[s];
	for(unsigned i = 0; i < n; i++){
		std::memcpy(h + b->in_off[i], normals[i], b->ch[i]->dim * sizeof(double));
		h[b->in_off[i] + b->ch[i]->dim] = uniforms[i];
	}
	// after each context's earlier work, and after the previous trajectory's records left
	for(unsigned i = 0; i < n; i++){
		EPHMC_HIP(hipEventRecord(b->ch_ev[i], b->ch[i]->stream));
		EPHMC_HIP(hipStreamWaitEvent(b->lead, b->ch_ev[i], 0));
	}
	const int rc = submit(*b, h, s);
	if(rc != EPHMC_OK){
		// an enqueue failed part way: wait for whatever did start, so that no batch work is left
		// running unordered with the contexts' own streams
		(void)hipStreamSynchronize(b->lead);
		for(auto &g : b->lanes) (void)hipStreamSynchronize(g.stream);
		return rc;
	}
	b->pending = true;
	return EPHMC_OK;
}

int ephmc_batch_sync(ephmc_batch *b, ephmc_traj_record *recs)
{
	if(!b){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(!b->pending){ set_error("batch: no trajectory in flight"); return EPHMC_ERR_ARG; }
	EPHMC_HIP(hipEventSynchronize(b->fin));
	b->pending = false;
	if(recs)
		for(size_t i = 0; i < b->ch.size(); i++){
			const double *s = b->h_rec + i * SCAL_COUNT;
			recs[i].accepted = s[SCAL_ACC] != 0;
			recs[i].boltzmann = s[SCAL_BOLTZ];
			recs[i].energy = s[SCAL_ENERGY];
			recs[i].logdet = s[SCAL_LOGDET_CHAIN];
			recs[i].dH = s[SCAL_DH];
		}
	return EPHMC_OK;
}

int ephmc_batch_info(const ephmc_batch *b, unsigned *groups, unsigned *launches)
{
	if(!b){ set_error("null argument"); return EPHMC_ERR_ARG; }
	unsigned l = 0;
	for(const auto &g : b->groups) l += (unsigned)g.seq.size();
	if(groups) *groups = (unsigned)b->groups.size();
	if(launches) *launches = l;
	return EPHMC_OK;
}

int ephmc_batch_destroy(ephmc_batch *b)
{
	release(b);
	return EPHMC_OK;
}

} // extern "C"
