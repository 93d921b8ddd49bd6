"""Synthetic ABC problem with a known truth, with the parameters of the
reference's integration script (test/test_part_conc_spt_abc.jl:23-77): 60x60x10
um box, 40x40 um field, 50 videos x 250 frames, dt 0.05 s, kmin 2, fixed depth
2 um, m = [1, 3] um^2/s, c = [1e8, 1e8] /mL, 200 DE bins up to 12. The
observed tracks come from the oracle at the truth (seed 2024)."""
import pathlib
import re

import numpy as np

import oracle_lib as O
from part_conc_spt_abc_amd.model import Model, proposal_record


def make_data(tmp, weighting="inverse-distance-squared", trials=12, samples=128, adaptive=True, gamma0=15.0,
              delta_gamma=0.02):
    model = Model(Lx=60.0, Ly=60.0, Lz=10.0, ax=40.0, ay=40.0, deltat=0.05, kmin=2, n_k=250, n_de=200, ub_de=12.0,
                  n_species=2)
    rng = np.random.default_rng(1)
    lam = 2e8 * model.volume_factor
    frames = [250] * 50
    parts = [int(x) for x in rng.poisson(lam, 50)]
    p = proposal_record(model, [1.0, 3.0], [1e8, 1e8], [2.0], 0xFFFFFFF0, 0, 50)
    K, DE = O.tracks_f32(O.system_from_abi(model.to_abi()), O.OracleProposal(p, frames, parts), 2024)
    data = tmp / "data.xml"
    data.write_text("<data>\n\t<ax>40.0</ax>\n\t<ay>40.0</ay>\n"
                    f"\t<number_of_frames>{','.join(map(str, frames))}</number_of_frames>\n"
                    "\t<deltat>0.05</deltat>\n"
                    f"\t<K>{','.join(str(int(k)) for k in K)}</K>\n"
                    f"\t<DE>{','.join('%.9g' % d for d in DE)}</DE>\n</data>\n")
    inp = tmp / "input.xml"
    inp.write_text(f"""<input>
	<data_file_path>{data}</data_file_path>
	<model>discrete-fixed-depth</model>
	<number_of_components>2</number_of_components>
	<weighting_scheme>{weighting}</weighting_scheme>
	<Lx>60.0</Lx>
	<Ly>60.0</Ly>
	<Lz>10.0</Lz>
	<kmin>2</kmin>
	<number_of_de_bins>200</number_of_de_bins>
	<ub_de>12.0</ub_de>
	<lb_m>0.25</lb_m>
	<ub_m>12.0</ub_m>
	<lb_c>2.5e7</lb_c>
	<ub_c>4.0e8</ub_c>
	<lb_az>0.5</lb_az>
	<ub_az>8.0</ub_az>
	<number_of_abc_samples>{samples}</number_of_abc_samples>
	<gamma_initial>{gamma0}</gamma_initial>
	<gamma_adaptive>{'true' if adaptive else 'false'}</gamma_adaptive>
	<delta_gamma>{delta_gamma}</delta_gamma>
	<ub_average_number_of_trials>{trials}</ub_average_number_of_trials>
	<output_file_path>{tmp / 'output.xml'}</output_file_path>
</input>
""")
    return inp, len(K), (model, frames, K, DE)


def read_output(path):
    txt = pathlib.Path(path).read_text()

    def vec(name):
        return np.array([float(x) for x in re.search(rf"<{name}>(.*?)</{name}>", txt, re.S).group(1).split(",")])

    def scal(name):
        return re.search(rf"<{name}>(.*?)</{name}>", txt, re.S).group(1).strip()

    return txt, vec, scal


