"""Write tests/golden/ibm64_records.json from the reference's frozen run records.

Sources (reference checkout, argument 1): protocol.json (generation seed, count, free vertices and the
per-sample prefixes/uniforms), run_r1/worker*/samples.jsonl (complex64 draws: A_index and CDF interval),
run_r1/worker*/audits.jsonl (complex128 block masses and full-vector witnesses of the audited prefixes),
run_r1/worker0/anchor.json (plan checksums), run_r1/analysis/verify_records.json (the run's KL-Chernoff
quality certificate), fullscore/analysis/scores.jsonl (complex128 amplitude at every drawn index).
Only data values are copied."""
import json
import sys
from pathlib import Path


def main(ref):
    ref = Path(ref)
    prot = json.loads((ref / 'protocol.json').read_text())
    g = prot['generation']
    samples = {}
    for w in range(8):
        for line in (ref / f'run_r1/worker{w}/samples.jsonl').read_text().splitlines():
            r = json.loads(line)
            d = r['draw']
            samples[r['sample_id']] = {
                'bits': r['input_data_bits'], 'u': r['draw_uniform'], 'A_index': d['A_index'],
                'lo': d['cdf_lower_numpy'], 'hi': d['cdf_upper_numpy'], 'margin': d['uniform_to_nearest_boundary'],
                'mass': d['scaled_mass_cdf'], 'weight': d['selected_scaled_weight'],
                'seconds': r['contraction']['seconds']}
    audits = {}
    for w in range(8):
        for line in (ref / f'run_r1/worker{w}/audits.jsonl').read_text().splitlines():
            r = json.loads(line)
            wit = r['witness']
            audits[r['audit_id']] = {'ref_mass': wit['reference_scaled_mass'],
                                     'approx_mass': wit['approximate_input_mass'],
                                     'witness': wit['uncapped_overlap_witness'], 'loss': wit['loss'],
                                     'overlap_abs': wit['normalized_conditional_overlap_abs'],
                                     'raw_approx_mass': r['raw_vector_witness']['approximate_input_mass'],
                                     'law_cdf_total': r['actual_cdf_law']['source_cdf_total'],
                                     'A_index': r['A_index'],
                                     'ref_selected_probability': r['reference_selected_probability'],
                                     'seconds': r['reference_contraction']['seconds'],
                                     'work_per_slice': r['reference_contraction']['work_per_slice'],
                                     'rmax': r['reference_contraction']['rmax']}
    scores = {}
    for line in (ref / 'fullscore/analysis/scores.jsonl').read_text().splitlines():
        r = json.loads(line)
        scores[r['sample_id']] = [r['A_index'], r['amplitude_real'], r['amplitude_imag'], r['probability_fp64']]
    anchor = json.loads((ref / 'run_r1/worker0/anchor.json').read_text())
    quality = json.loads((ref / 'run_r1/analysis/verify_records.json').read_text())['quality']
    out = {
        'generation': {'seed': g['seed'], 'count': g['count'], 'n': g['n'], 'free_vertices': g['free_vertices'],
                       'mixture_weight': g['mixture_weight'],
                       'first_records': [{k: rec[k] for k in ('sample_id', 'branch_uniform', 'draw_uniform',
                                                                 'input_data_bits')} for rec in g['records'][:16]]},
        'anchor': {'bits': anchor['input_data_bits'], 'rmax': anchor['contraction']['rmax'],
                   'terminal_rank': anchor['contraction']['terminal_rank'],
                   'work_per_slice': anchor['contraction']['work_per_slice'],
                   'slices_evaluated': anchor['contraction']['slices_evaluated']},
        'samples': {str(k): v for k, v in sorted(samples.items())},
        'audits': {str(k): v for k, v in sorted(audits.items())},
        'scores': {str(k): v for k, v in sorted(scores.items())},
        'certificate': {k: quality[k] for k in ('audit_count', 'alpha', 'observed_mean_loss',
                                                 'population_mean_loss_upper', 'overlap_lower',
                                                 'state_fidelity_lower')},
    }
    dst = Path(__file__).resolve().parent / 'ibm64_records.json'
    dst.write_text(json.dumps(out, separators=(',', ':')))
    print(dst, dst.stat().st_size, len(samples), len(audits), len(scores))


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
